"""GPU factor and solve on a reordered supernode numbering (-refine SET*,
TSPB*):
the HIP path against the oracle on the same refined structure, and against
the unrefined run of the same system.  Tolerances as in the other parity
tests: panels and solutions within 1e-10 relative of the oracle's."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


def _solve(gen, ordering, refine, ldl, nrhs=3, seed=0):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    opts = sp.symPACKOptions(ordering=ordering, decomposition="LDL" if ldl else "LL",
                             order_refinement_str=refine)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(opts)
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    print("case n=%d ordering=%s refine=%s ldl=%s" % (n, ordering, refine, ldl), flush=True)
    Mx.Factorize()
    xt = np.random.default_rng(seed).uniform(-1, 1, (n, nrhs))
    B = A.matvec(xt)
    Mx.Solve(B, nrhs)
    X = Mx.GetSolution(nrhs).reshape(n, nrhs)
    return A, Mx, X, B


@pytest.mark.parametrize("name,gen,ordering,code,ldl", [
    ("st27_8_mmd_set", lambda: M.stencil_3d_27pt(8, seed=2), "MMD", "SET", False),
    ("diff3d_14_amd_set21", lambda: M.diffusion_3d_7pt(14, seed=3), "AMD", "SET21", False),
    ("lap2d_48_mmd_set10", lambda: M.laplacian_2d_5pt(48), "MMD", "SET10", False),
    ("diff3d_12_mmd_set_ldl", lambda: M.diffusion_3d_7pt(12, seed=5), "MMD", "SET", True),
    ("st27_8_mmd_tspb", lambda: M.stencil_3d_27pt(8, seed=2), "MMD", "TSPB", False),
    ("lap2d_48_amd_tspb1", lambda: M.laplacian_2d_5pt(48), "AMD", "TSPB1", False),
    ("diff3d_12_mmd_tspb_ldl", lambda: M.diffusion_3d_7pt(12, seed=5), "MMD", "TSPB", True),
])
def test_refined_factor_and_solve_match_oracle(name, gen, ordering, code, ldl):
    A, Mx, X, B = _solve(gen, ordering, code, ldl)
    S = onum.Structure.of(Mx)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    Xo = onum.solve(S, F, B, ldl)
    if not ldl:
        op = onum.panels(S, F)
        num = den = 0.0
        for I in range(1, S.nsuper + 1):
            g = Mx.GetFactorPanel(I)
            o = op[I - 1]
            w = g.shape[1]
            mask = np.ones_like(o, bool)
            mask[:w, :w] = np.tril(np.ones((w, w), bool))
            num += np.sum((g[mask] - o[mask]) ** 2)
            den += np.sum(o[mask] ** 2)
        assert np.sqrt(num / den) < 1e-10
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10
    # the same system without the reordering
    _, _, X0, _ = _solve(gen, ordering, "NO", ldl)
    assert np.linalg.norm(X - X0) / np.linalg.norm(X0) < 1e-10
