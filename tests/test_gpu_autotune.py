"""The Schur-launch autotuner: on the first factorization of a large enough
structure every etree level's Schur launch is timed under each Schur kernel
(64 x 64 tiles staged through LDS, 128 x 64 tiles, 64 x 64 tiles with one
wave per tile), task order and k split and the fastest is kept.  The tuned factorization must give the
oracle's solution (1e-10 relative, the parity bar), and so must the
refactorizations that replay the recaptured launch sequence; the tuner's
record must be consistent (one record per Schur launch, the chosen time no
worse than the default's, configurations within the candidate set)."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


def _factor(A, autotune, ldl=False, precision="FP64"):
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL" if ldl else "LL", relax=rp,
                             autotune=autotune, precision=precision))
    m.SymbolicFactorization(A)
    m.DistributeMatrix(A)
    m.Factorize()
    return m


@pytest.mark.parametrize("ldl", [False, True], ids=["LL", "LDL"])
def test_autotuned_factor_matches_oracle(ldl):
    gen = (lambda: M.shifted_laplacian_3d_7pt(32, 0.3)) if ldl else (lambda: M.diffusion_3d_7pt(32, seed=3))
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    print("case 32^3 ldl=%s autotune" % ldl, flush=True)
    m = _factor(A, True, ldl)
    log = m.autotune_log()
    assert m.stats()["schur_flops"] >= 1e9
    assert len(log) > 0
    for r in log:
        assert 1 <= r["split"] <= 8 and r["order"] in (0, 1, 2, 3, 4, 5)
        assert 0 < r["ms_best"] <= r["ms_default"] * (1 + 1e-12)
    rng = np.random.default_rng(7)
    B = rng.uniform(-1, 1, (n, 3))
    S = onum.Structure.of(m)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    Xo = onum.solve(S, F, B, ldl)
    for rep in range(3):  # the tuned sequence, then replays of its captured graph
        if rep:
            m.Factorize()
        m.Solve(B, 3)
        X = m.GetSolution(3).reshape(n, 3)
        assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10, rep
    m0 = _factor(A, False, ldl)
    assert m0.autotune_log() == []
    m0.Solve(B, 3)
    X0 = m0.GetSolution(3).reshape(n, 3)
    assert np.linalg.norm(X0 - Xo) / np.linalg.norm(Xo) < 1e-10


def test_autotuned_fp32_ir_converges():
    n, cp, ri, v = M.stencil_3d_27pt(28, seed=5)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    print("case 27-point 28^3 FP32_IR autotune", flush=True)
    m = _factor(A, True, precision="FP32_IR")
    assert len(m.autotune_log()) > 0
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, 4))
    m.Solve(B, 4)
    X = m.GetSolution(4).reshape(n, 4)
    st = m.stats()
    assert st["ir_converged"] == 1
    assert onum.backward_error(A, X, B) <= 1e-12


def test_small_structures_are_not_tuned():
    n, cp, ri, v = M.diffusion_3d_7pt(10, seed=1)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    m = _factor(A, True)
    assert m.stats()["schur_flops"] < 1e9
    assert m.autotune_log() == []


# Each Schur kernel forced on every level (schur_kernel 2: 128 x 64 tiles,
# k_schur_wide; 3: 64 x 64 tiles with one wave per tile, k_schur_w1): plain
# updates, merged chain updates (merge_updates=2 makes every qualifying chain
# child defer into its parent) and deferred column cuts, LL^T and static
# LDL^T, FP64 and FP32 + refinement.  Parity bar: solutions within 1e-10
# relative of the oracle's.
FORCED_CASES = [
    ("diff3d_16_b24_merged", lambda: M.diffusion_3d_7pt(16, seed=4), 24, 2),
    ("st27_12_b20_merged", lambda: M.stencil_3d_27pt(12, seed=3), 20, 2),
    ("diff3d_24_b150", lambda: M.diffusion_3d_7pt(24, seed=8), 150, 1),
    ("lap2d_64_unlimited", lambda: M.laplacian_2d_5pt(64), 0, 1),
]


@pytest.mark.parametrize("kernel", [2, 3], ids=["wide", "w1"])
@pytest.mark.parametrize("ldl", [False, True], ids=["LL", "LDL"])
@pytest.mark.parametrize("name,gen,maxsize,merge", FORCED_CASES, ids=[c[0] for c in FORCED_CASES])
def test_forced_schur_kernel_matches_oracle(name, gen, maxsize, merge, ldl, kernel):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    if maxsize:
        rp.SetMaxSize(maxsize)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="NDBOX" if maxsize else "MMD", decomposition="LDL" if ldl else "LL",
                             relax=rp, merge_updates=merge, schur_kernel=kernel))
    m.SymbolicFactorization(A)
    m.DistributeMatrix(A)
    print("case %s ldl=%s schur_kernel=%d" % (name, ldl, kernel), flush=True)
    m.Factorize()
    rng = np.random.default_rng(9)
    B = rng.uniform(-1, 1, (n, 2))
    m.Solve(B, 2)
    X = m.GetSolution(2).reshape(n, 2)
    S = onum.Structure.of(m)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    Xo = onum.solve(S, F, B, ldl)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10


@pytest.mark.parametrize("kernel", [2, 3], ids=["wide", "w1"])
def test_forced_schur_kernel_fp32_ir(kernel):
    n, cp, ri, v = M.stencil_3d_27pt(16, seed=2)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision="FP32_IR", schur_kernel=kernel))
    m.SymbolicFactorization(A)
    m.DistributeMatrix(A)
    print("case 27-point 16^3 FP32_IR schur_kernel=%d" % kernel, flush=True)
    m.Factorize()
    B = np.random.default_rng(4).standard_normal((n, 3))
    m.Solve(B, 3)
    X = m.GetSolution(3).reshape(n, 3)
    assert m.stats()["ir_converged"] == 1
    assert onum.backward_error(A, X, B) <= 1e-12
