"""GPU: the two block-factor paths and the split Schur update in one
factorization.  Etree levels with at most 256 TRSM row tiles run the fused
block factor + panel TRSM (one workgroup per row tile, each factoring the
diagonal block again; the last to load it writes it back), wider levels run
the block factor and the TRSM as two launches, and Schur launches with fewer
than 256 tiles split each tile's k chunks over several workgroups.  The
launch counters show that both block-factor paths ran; panels (LL, 1e-10
relative, Frobenius) and solutions (1e-10 relative) must match the oracle."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu

# 24^3 and 28^3 (width <= 150) each have one etree level above 256 row tiles
# (433 and 991); every other level is fused
CASES = [
    ("diff3d_24_b150", lambda: M.diffusion_3d_7pt(24, seed=41), 150, False),
    ("diff3d_24_b150_ldl", lambda: M.diffusion_3d_7pt(24, seed=42), 150, True),
    ("diff3d_28_b150", lambda: M.diffusion_3d_7pt(28, seed=43), 150, False),
]


@pytest.mark.parametrize("name,gen,bmax,ldl", CASES, ids=[c[0] for c in CASES])
def test_fused_and_split_paths_match_oracle(name, gen, bmax, ldl):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(bmax)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL" if ldl else "LL", relax=rp))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    print("case", name, "n", n, "nsuper", Mx.TotalSupernodeCnt(), flush=True)
    Mx.set_profile(True)
    Mx.reset_kernel_stats()
    Mx.Factorize()
    Mx.set_profile(False)
    launches = {k: Mx.kernel_stats(k)[0] for k in (0, 7, 9, 2)}
    print("launches", launches, flush=True)
    assert launches[9] > 0   # fused block factor + TRSM
    assert launches[0] > 0   # separate block factor (wide levels)
    assert launches[2] > 0
    S = onum.Structure.of(Mx)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
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
    B = np.random.default_rng(44).uniform(-1, 1, (n, 2))
    Mx.Solve(B, 2)
    X = Mx.GetSolution(2).reshape(n, 2)
    Xo = onum.solve(S, F, B, ldl).reshape(n, 2)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10
