"""GPU parity: HIP factor and solve against the oracle (oracle/supernodal.c)
on the same matrices, orderings and symbolic structure.

Tolerances: factor panels agree to 1e-10 relative (Frobenius, FP64 with a
different summation order); solutions within 1e-10 relative of the oracle's;
backward error within 10x of the oracle's (the drop-in bar).
"""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


def _run(gen, ordering, ldl=False, nrhs=1, relax_max=None, seed=0):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    if relax_max is not None:
        rp.SetMaxSize(relax_max)
    opts = sp.symPACKOptions(ordering=ordering, decomposition="LDL" if ldl else "LL", relax=rp)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(opts)
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    print("case n=%d ordering=%s ldl=%s nrhs=%d nsuper=%d" % (n, ordering, ldl, nrhs,
                                                             Mx.TotalSupernodeCnt()), flush=True)
    Mx.Factorize()
    rng = np.random.default_rng(seed)
    xt = rng.uniform(-1, 1, (n, nrhs))
    B = A.matvec(xt)
    Mx.Solve(B, nrhs)
    X = Mx.GetSolution(nrhs).reshape(n, nrhs)
    S = onum.Structure.of(Mx)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    Xo = onum.solve(S, F, B, ldl)
    return A, Mx, S, F, X, Xo, B, xt


CASES = [
    ("lap2d_64_mmd", lambda: M.laplacian_2d_5pt(64), "MMD", None),
    ("lap2d_40_ndgrid", lambda: M.laplacian_2d_5pt(40), "NDGRID", None),
    ("diff3d_12_ndbox_b150", lambda: M.diffusion_3d_7pt(12, seed=3), "NDBOX", 150),
    ("diff3d_16_ndbox_b150", lambda: M.diffusion_3d_7pt(16, seed=4), "NDBOX", 150),
    ("diff3d_10_ndbox_unlimited", lambda: M.diffusion_3d_7pt(10, seed=5), "NDBOX", None),
    ("st27_10_mmd", lambda: M.stencil_3d_27pt(10, seed=6), "MMD", None),
    ("rand_spd_600_rcm", lambda: M.random_spd(600, 0.01, 7), "RCM", None),
]


@pytest.mark.parametrize("name,gen,ordering,bmax", CASES, ids=[c[0] for c in CASES])
def test_factor_panels_and_solve_ll(name, gen, ordering, bmax):
    A, Mx, S, F, X, Xo, B, xt = _run(gen, ordering, relax_max=bmax, nrhs=2)
    op = onum.panels(S, F)
    num = den = 0.0
    for I in range(1, S.nsuper + 1):
        g = Mx.GetFactorPanel(I)
        o = op[I - 1]
        w = g.shape[1]
        # the strict upper part of the diagonal block is not part of L
        mask = np.ones_like(o, bool)
        mask[:w, :w] = np.tril(np.ones((w, w), bool))
        num += np.sum((g[mask] - o[mask]) ** 2)
        den += np.sum(o[mask] ** 2)
    assert np.sqrt(num / den) < 1e-10
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10
    be_g = onum.backward_error(A, X, B)
    be_o = onum.backward_error(A, Xo, B)
    assert be_g <= 10 * max(be_o, 1e-16)


@pytest.mark.parametrize("name,gen,ordering,bmax", CASES[:5], ids=[c[0] for c in CASES[:5]])
def test_ldl_matches_oracle(name, gen, ordering, bmax):
    A, Mx, S, F, X, Xo, B, xt = _run(gen, ordering, ldl=True, relax_max=bmax, nrhs=3)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10
    assert onum.backward_error(A, X, B) <= 10 * max(onum.backward_error(A, Xo, B), 1e-16)


def test_indefinite_ldl_shifted_laplacian():
    A, Mx, S, F, X, Xo, B, xt = _run(lambda: M.shifted_laplacian_3d_7pt(12, 0.3), "NDBOX",
                                     ldl=True, relax_max=150, nrhs=4)
    # static pivots on an indefinite matrix: same bar as the drop-in, relative
    # to the oracle's own backward error
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-9
    assert onum.backward_error(A, X, B) <= 10 * max(onum.backward_error(A, Xo, B), 1e-16)


@pytest.mark.parametrize("nrhs", [1, 5, 16, 64, 70])
def test_many_rhs(nrhs):
    A, Mx, S, F, X, Xo, B, xt = _run(lambda: M.diffusion_3d_7pt(10, seed=9), "NDBOX",
                                     relax_max=150, nrhs=nrhs)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-10


def test_not_spd_raises():
    n, cp, ri, v = M.laplacian_2d_5pt(16)
    v = v.copy()
    v[cp[100]] = -5.0  # make one diagonal entry negative
    A = sp.DistSparseMatrix(n, cp, ri, v)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="MMD"))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    with pytest.raises(sp.NotHPDError) as e:
        Mx.Factorize()
    info, F = onum.factor(onum.Structure.of(Mx), A)
    assert e.value.column == info > 0


def test_fp32_ir_reaches_fp64_accuracy():
    n, cp, ri, v = M.stencil_3d_27pt(12, seed=11)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision="FP32_IR", ir_tol=1e-12))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    rng = np.random.default_rng(5)
    B = rng.standard_normal((n, 4))
    Mx.Solve(B, 4)
    X = Mx.GetSolution(4)
    st = Mx.stats()
    assert st["backward_error"] <= 1e-12
    assert onum.backward_error(A, X, B) <= 1e-12


def test_fp32_ir_ldl_indefinite():
    n, cp, ri, v = M.shifted_laplacian_3d_7pt(12, 0.3)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL", relax=rp, precision="FP32_IR",
                              ir_tol=1e-12))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    rng = np.random.default_rng(6)
    B = rng.uniform(-1, 1, (n, 3))
    Mx.Solve(B, 3)
    X = Mx.GetSolution(3)
    assert Mx.stats()["backward_error"] <= 1e-12
    assert onum.backward_error(A, X, B) <= 1e-12


def _panels(Mx, S):
    return [Mx.GetFactorPanel(I) for I in range(1, S.nsuper + 1)]


def _rel(a, b):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(y ** 2)) for y in b)
    return np.sqrt(num / den)


@pytest.mark.parametrize("prec", ["FP64", "FP32_IR"])
def test_graph_replays_match_eager(prec):
    n, cp, ri, v = M.diffusion_3d_7pt(12, seed=21)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)

    def make(graphs):
        Mx = sp.symPACKMatrix2D()
        Mx.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision=prec))
        Mx.set_graphs(graphs)
        Mx.SymbolicFactorization(A)
        Mx.DistributeMatrix(A)
        return Mx

    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (n, 2))
    E = make(False)
    E.Factorize()
    S = onum.Structure.of(E)
    P0 = _panels(E, S)
    E.Solve(B, 2)
    X0 = E.GetSolution(2)
    G = make(True)
    tol = 1e-12 if prec == "FP64" else 1e-5
    for rep in range(3):
        print("graph replay", prec, rep, flush=True)
        G.Factorize()
        assert _rel(_panels(G, S), P0) < tol
        G.Solve(B, 2)
        X = G.GetSolution(2)
        assert np.linalg.norm(X - X0) / np.linalg.norm(X0) < 1e-10
        # eager work and a synchronization between replays
        E.Solve(B[:, :1], 1)
        E.GetSolution(1)
    # the launch counters see every replayed kernel
    G.reset_kernel_stats()
    E.reset_kernel_stats()
    G.Factorize()
    E.Factorize()
    for k in range(10):
        assert G.kernel_stats(k)[0] == E.kernel_stats(k)[0]


def test_profile_mode_records_kernel_times():
    n, cp, ri, v = M.diffusion_3d_7pt(10, seed=22)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    Mx.reset_kernel_stats()
    Mx.set_profile(True, kinds=[2])
    for _ in range(3):
        Mx.Factorize()
    Mx.set_profile(False)
    launches, ms = Mx.kernel_stats(2)
    assert launches > 0 and ms > 0
    kinds, grids, times = Mx.launch_trace()
    assert len(kinds) == launches and set(kinds) == {2} and np.all(times > 0)


def test_fp32_pcg_ill_conditioned_27pt():
    # diagonal = row abs-sum * (1 + 1e-6): kappa ~ 1e6-1e7, beyond plain
    # fixed-point refinement with an FP32 factor
    n, cp, ri, v = M.stencil_3d_27pt(24, seed=13)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision="FP32_IR", ir_tol=1e-12))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    rng = np.random.default_rng(14)
    B = rng.standard_normal((n, 16))
    Mx.Solve(B, 16)
    X = Mx.GetSolution(16)
    st = Mx.stats()
    print("pcg iterations", st["ir_iterations"], "berr", st["backward_error"], flush=True)
    assert st["backward_error"] <= 1e-12
    assert onum.backward_error(A, X, B) <= 1e-12


def test_fp32_nearly_singular_spd():
    # graph Laplacian + 1e-10 I: SPD in FP64, at the edge of FP32
    n, cp, ri, v = M.diffusion_3d_7pt(16, seed=15, shift=1e-10)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision="FP32_IR", ir_tol=1e-12,
                              ir_max_iter=300))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    rng = np.random.default_rng(16)
    B = rng.standard_normal((n, 2))
    Mx.Solve(B, 2)
    X = Mx.GetSolution(2)
    st = Mx.stats()
    print("shift", st["fp32_shift"], "iterations", st["ir_iterations"], "berr", st["backward_error"], flush=True)
    assert onum.backward_error(A, X, B) <= 1e-12


def test_ldl_48cube_refactorizations_match_oracle():
    # regression: a read-after-write race on the LDS diagonal tile of the
    # diagonal-block factor showed up here as backward errors ~1e-2 that
    # changed from one factorization to the next
    n, cp, ri, v = M.shifted_laplacian_3d_7pt(48, 0.3)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL", relax=rp))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    rng = np.random.default_rng(17)
    B = rng.uniform(-1, 1, (n, 2))
    S = onum.Structure.of(Mx)
    info, F = onum.factor(S, A, True)
    be_o = onum.backward_error(A, onum.solve(S, F, B, True), B)
    for rep in range(3):
        Mx.Factorize()
        Mx.Solve(B, 2)
        be = onum.backward_error(A, Mx.GetSolution(2), B)
        print("refactor", rep, "berr %.2e oracle %.2e" % (be, be_o), flush=True)
        assert be <= 10 * max(be_o, 1e-16)


def test_disabled_dependency_driven_solve_is_rejected():
    """solve_dataflow=1 names the dependency-driven sweeps, which are switched
    off in this build: Solve raises before launching anything."""
    n, cp, ri, v = M.laplacian_2d_5pt(16)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="MMD", solve_dataflow=True))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    with pytest.raises(ValueError):
        Mx.Solve(np.ones(n), 1)
