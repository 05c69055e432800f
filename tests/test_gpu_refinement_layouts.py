"""FP32 factor + FP64 refinement with n not a multiple of 64 and more than 64
right-hand sides: the right-hand side gather, the scaled solution scatter
with accumulation and the multi-pass SpMV all take their partial last tile
and a second block of right-hand sides.  Bar: backward error <= 1e-12 (the
refinement target) and the solution within 1e-8 relative of the FP64
oracle's."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,gen,ldl", [
    ("diff3d_11_ll", lambda: M.diffusion_3d_7pt(11, seed=7), False),
    ("shifted_lap_11_ldl", lambda: M.shifted_laplacian_3d_7pt(11, 0.3), True),
])
def test_fp32_ir_partial_tiles_many_rhs(name, gen, ldl):
    n, cp, ri, v = gen()
    assert n % 64 != 0
    nrhs = 70
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL" if ldl else "LL", relax=rp,
                              precision="FP32_IR", ir_tol=1e-12))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    print("case %s n=%d nrhs=%d" % (name, n, nrhs), flush=True)
    Mx.Factorize()
    B = np.random.default_rng(3).uniform(-1, 1, (n, nrhs))
    Mx.Solve(B, nrhs)
    X = Mx.GetSolution(nrhs).reshape(n, nrhs)
    assert Mx.stats()["backward_error"] <= 1e-12
    assert onum.backward_error(A, X, B) <= 1e-12
    Xo = onum.factor_and_solve(Mx, A, B, ldl)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-8


def test_refinement_reports_convergence():
    n, cp, ri, v = M.shifted_laplacian_3d_7pt(10, 0.3)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    B = np.random.default_rng(4).uniform(-1, 1, (n, 2))
    for maxit, tol, want in [(30, 1e-12, 1), (1, 1e-30, 0)]:
        Mx = sp.symPACKMatrix2D()
        Mx.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL", precision="FP32_IR",
                                  ir_tol=tol, ir_max_iter=maxit))
        Mx.SymbolicFactorization(A)
        Mx.DistributeMatrix(A)
        Mx.Factorize()
        Mx.Solve(B, 2)
        st = Mx.stats()
        assert st["ir_converged"] == want
        if want:
            assert st["backward_error"] <= tol
        else:
            assert st["ir_iterations"] == maxit and st["backward_error"] > tol
