"""GPU parity of merged Schur updates: with narrow supernodes every separator
becomes a chain, so many supernodes defer the part of their update outside
their parent's columns into the parent's update.  The factor and solution
must match the oracle (which applies every update from its own source, as
the reference does) to the same bars as the plain path: panels within 1e-10
relative (Frobenius), solutions within 1e-10 relative, backward error within
10x of the oracle's; FP32+IR reaches 1e-12."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu

CASES = [
    ("diff3d_16_b24", lambda: M.diffusion_3d_7pt(16, seed=4), 24),
    ("st27_12_b20", lambda: M.stencil_3d_27pt(12, seed=3), 20),
    ("diff3d_20_b48", lambda: M.diffusion_3d_7pt(20, seed=8), 48),
]


def _factor(gen, maxsize, ldl=False, merge=2, precision="FP64"):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(maxsize)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL" if ldl else "LL", relax=rp,
                              merge_updates=merge, precision=precision))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    nm = int(np.sum(Mx.GetUpdateMerges() >= 0))
    print("case n=%d maxsize=%d ldl=%s merge=%s merges=%d prec=%s" % (n, maxsize, ldl, merge, nm, precision),
          flush=True)
    Mx.Factorize()
    return A, Mx, nm


@pytest.mark.parametrize("ldl", [False, True], ids=["LL", "LDL"])
@pytest.mark.parametrize("name,gen,maxsize", CASES, ids=[c[0] for c in CASES])
def test_merged_updates_match_oracle(name, gen, maxsize, ldl):
    A, Mx, nm = _factor(gen, maxsize, ldl=ldl)
    assert nm > 0
    n = A.size
    rng = np.random.default_rng(2)
    xt = rng.uniform(-1, 1, (n, 3))
    B = A.matvec(xt)
    Mx.Solve(B, 3)
    X = Mx.GetSolution(3).reshape(n, 3)
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
    assert onum.backward_error(A, X, B) <= 10 * max(onum.backward_error(A, Xo, B), 1e-16)
    # the plain (unmerged) path gives the same solution
    A2, M2, nm2 = _factor(gen, maxsize, ldl=ldl, merge=0)
    assert nm2 == 0
    M2.Solve(B, 3)
    X2 = M2.GetSolution(3).reshape(n, 3)
    assert np.linalg.norm(X - X2) / np.linalg.norm(X2) < 1e-11


def test_merged_updates_fp32_ir():
    A, Mx, nm = _factor(lambda: M.stencil_3d_27pt(12, seed=3), 20, precision="FP32_IR")
    assert nm > 0
    B = np.random.default_rng(6).standard_normal((A.size, 4))
    Mx.Solve(B, 4)
    X = Mx.GetSolution(4)
    assert Mx.stats()["backward_error"] <= 1e-12
    assert onum.backward_error(A, X, B) <= 1e-12
