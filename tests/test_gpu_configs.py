"""GPU tests of the precision and serving configurations at their own shapes.

* 27-point SPD stencil (diagonal = row abs-sum * (1 + 1e-6), kappa ~1e6-1e8),
  NDBOX, 16 Gaussian right-hand sides, FP32 factor + FP64 refinement:
  against the oracle's FP64 factor + solve at 16^3 and 20^3 (relative
  difference <= 1e-8: both solutions are backward stable, so they may differ
  by up to kappa * u; backward error <= 1e-12; the refinement reports
  convergence), and at the full 60^3 size by the properties alone
  (backward error <= 1e-12 computed here on the host, convergence).
* The serving stream's shapes: 2D 5-point / 3D 7-point random-coefficient
  Laplacians with n log-uniform in [2e3, 2e5], follow-up solves with 1-32
  right-hand sides, FP64 and FP32+IR: every result's backward error, and the
  same result alone (one slot, one request at a time) and in the mix (four
  slots, everything submitted at once).
"""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from sympack_amd.serving import SolverService, make_request_stream
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


def _fp32_ir(A, nrhs, seed):
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision="FP32_IR", ir_tol=1e-12))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    B = np.random.default_rng(seed).standard_normal((A.size, nrhs))
    Mx.Solve(B, nrhs)
    X = Mx.GetSolution(nrhs).reshape(A.size, nrhs)
    return Mx, B, X


@pytest.mark.parametrize("k", [16, 20])
def test_27pt_fp32_ir_against_oracle(k):
    n, cp, ri, v = M.stencil_3d_27pt(k, seed=k)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    print("27-point %d^3 FP32+IR nrhs=16" % k, flush=True)
    Mx, B, X = _fp32_ir(A, 16, seed=100 + k)
    st = Mx.stats()
    Xo = onum.factor_and_solve(Mx, A, B)
    rel = np.linalg.norm(X - Xo) / np.linalg.norm(Xo)
    be = onum.backward_error(A, X, B)
    print("steps %d converged %d rel.diff vs oracle %.2e berr %.2e (oracle %.2e)" %
          (st["ir_iterations"], st["ir_converged"], rel, be, onum.backward_error(A, Xo, B)), flush=True)
    assert st["ir_converged"] == 1
    assert be <= 1e-12
    assert rel <= 1e-8


def test_27pt_60cube_fp32_ir_full_size():
    n, cp, ri, v = M.stencil_3d_27pt(60, seed=0)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    print("27-point 60^3 FP32+IR nrhs=16", flush=True)
    Mx, B, X = _fp32_ir(A, 16, seed=1)
    st = Mx.stats()
    be = onum.backward_error(A, X, B)
    print("steps %d converged %d berr %.2e (library %.2e)" %
          (st["ir_iterations"], st["ir_converged"], be, st["backward_error"]), flush=True)
    assert st["ir_converged"] == 1
    assert be <= 1e-12


def _stream(n_requests, seed):
    stream = make_request_stream(n_requests, 1000.0, seed=seed)
    mats, rhs = {}, []
    for t, kind, key, p in stream:
        if kind == "factor":
            dim, k, s = p
            n, cp, ri, v = M.random_laplacian("2d" if dim == 2 else "3d", k, s)
            mats[key] = (dim, sp.DistSparseMatrix(n, cp, ri, v))
        else:
            rhs.append((key, np.random.default_rng(p[1]).standard_normal((mats[key][1].size, p[0]))))
    return stream, mats, rhs


def _opts(dim, prec):
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    return sp.symPACKOptions(ordering="NDGRID" if dim == 2 else "NDBOX", relax=rp, precision=prec)


def _run_stream(stream, mats, rhs, prec, slots):
    out = {}
    with SolverService(slots=slots, max_batch_cols=64) as s:
        futs, j = [], 0
        for t, kind, key, p in stream:
            if kind == "factor":
                futs.append(("f", key, s.submit_factor(key, mats[key][1], _opts(mats[key][0], prec))))
                if slots == 1:
                    futs[-1][2].result(timeout=300)
            else:
                futs.append(("s", j, s.submit_solve(key, rhs[j][1])))
                if slots == 1:
                    futs[-1][2].result(timeout=300)
                j += 1
        for kind, j, f in futs:
            r = f.result(timeout=300)
            if kind == "s":
                out[j] = r
        rep = s.report()
    return out, rep


@pytest.mark.parametrize("prec", ["FP64", "FP32_IR"])
def test_serving_stream_shapes_alone_equal_mix(prec):
    stream, mats, rhs = _stream(40, seed=11)
    sizes = sorted(A.size for _, A in mats.values())
    print("stream: %d factors n in [%d, %d], %d solves, %s" % (len(mats), sizes[0], sizes[-1], len(rhs), prec),
          flush=True)
    assert sizes[-1] >= 2e4  # the stream reaches the upper size decade
    alone, _ = _run_stream(stream, mats, rhs, prec, slots=1)
    print("alone done", flush=True)
    mix, rep = _run_stream(stream, mats, rhs, prec, slots=4)
    print("mix done: solve p50 %.2f ms p99 %.2f ms" % (rep["solve"]["p50_ms"], rep["solve"]["p99_ms"]), flush=True)
    assert rep["solve"]["count"] == len(rhs)
    for j, (key, B) in enumerate(rhs):
        A = mats[key][1]
        a, b = alone[j], mix[j]
        be_a, be_b = onum.backward_error(A, a, B), onum.backward_error(A, b, B)
        rel = np.linalg.norm(a - b) / np.linalg.norm(a)
        if prec == "FP64":
            assert be_a <= 1e-14 and be_b <= 1e-14
            assert rel <= 1e-11
        else:
            assert be_a <= 1e-12 and be_b <= 1e-12
            assert rel <= 1e-8


def test_ldl_80cube_fp32_gmres_ir_converges():
    """Static LDL^T of the shifted 7-point Laplacian L_h - 0.3 I on 80^3
    (indefinite), FP32 factor + FP64 GMRES refinement, 64 right-hand sides:
    backward error <= 1e-12 (computed here on the host) and the refinement
    reports convergence."""
    n, cp, ri, v = M.shifted_laplacian_3d_7pt(80, 0.3)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL", relax=rp, precision="FP32_IR",
                              ir_tol=1e-12))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    print("80^3 shifted Laplacian LDL^T FP32 + GMRES, nrhs=64", flush=True)
    B = np.random.default_rng(2).uniform(-1, 1, (n, 64))
    for rep in range(2):
        Mx.Factorize()
        Mx.Solve(B, 64)
        X = Mx.GetSolution(64).reshape(n, 64)
        st = Mx.stats()
        F = A.to_scipy_full()
        R = B - F @ X
        normA = abs(F).sum(axis=1).max()
        be = float(np.max(np.abs(R).max(axis=0) / (normA * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))))
        print("factorization %d: steps %d converged %d berr %.2e (library %.2e)" %
              (rep, st["ir_iterations"], st["ir_converged"], be, st["backward_error"]), flush=True)
        assert st["ir_converged"] == 1
        assert be <= 1e-12
