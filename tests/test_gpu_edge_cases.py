"""GPU factor + solve on degenerate shapes, against exact answers and the
oracle (oracle/numeric.py): a 1x1 matrix, a diagonal matrix (supernodes
without off-diagonal rows, so no Schur update at all) and a block-diagonal
matrix of two unconnected Laplacians (an elimination forest with two roots).
LL^T and static LDL^T (the LDL^T diagonal case has pivots of both signs),
FP64 and FP32 factor + FP64 refinement.

Tolerances (FP64 / FP32 + IR): solutions within 1e-12 / 1e-10 relative of
the exact solution of the diagonal systems and within 1e-10 / 1e-7 relative
of the oracle's FP64 solution otherwise (a different summation order or a
refinement stopped at backward error 1e-12, times the condition number);
backward error <= 1e-14 / 1e-12 with the refinement converged.
"""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from sympack_amd.matrices import _lower_csc
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


def _solve(n, cp, ri, v, ordering, ldl, nrhs=3, seed=0, prec="FP64"):
    A = sp.DistSparseMatrix(n, cp, ri, v)
    M_ = sp.symPACKMatrix2D()
    M_.Init(sp.symPACKOptions(ordering=ordering, decomposition="LDL" if ldl else "LL", precision=prec,
                              ir_tol=1e-12))
    M_.SymbolicFactorization(A)
    M_.DistributeMatrix(A)
    print("case n=%d ordering=%s ldl=%s precision=%s nsuper=%d" % (n, ordering, ldl, prec, M_.TotalSupernodeCnt()),
          flush=True)
    M_.Factorize()
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, (n, nrhs))
    M_.Solve(B, nrhs)
    X = M_.GetSolution(nrhs)  # n x nrhs
    if prec == "FP32_IR":
        st = M_.stats()
        assert st["ir_converged"] and st["backward_error"] <= 1e-12
    return A, M_, B, X


PRECS = ["FP64", "FP32_IR"]


def _bwd_tol(prec):
    return 1e-14 if prec == "FP64" else 1e-12


def _fwd_tol(prec, exact):
    if exact:
        return 1e-12 if prec == "FP64" else 1e-10
    return 1e-10 if prec == "FP64" else 1e-7


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ldl", [False, True])
def test_one_by_one(ldl, prec):
    d = -4.0 if ldl else 4.0
    n, cp, ri, v = _lower_csc(1, [0], [0], [d])
    A, _, B, X = _solve(n, cp, ri, v, "NATURAL", ldl, prec=prec)
    assert np.allclose(X, B / d, rtol=_fwd_tol(prec, True), atol=0)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ldl", [False, True])
def test_diagonal_matrix(ldl, prec):
    n = 300
    rng = np.random.default_rng(1)
    d = rng.uniform(1.0, 10.0, n)
    if ldl:
        d[::3] *= -1.0  # indefinite: static LDL^T accepts any non-zero pivot
    i = np.arange(n)
    n, cp, ri, v = _lower_csc(n, i, i, d)
    A, M_, B, X = _solve(n, cp, ri, v, "MMD", ldl, prec=prec)
    assert np.linalg.norm(X - B / d[:, None]) <= _fwd_tol(prec, True) * np.linalg.norm(B / d[:, None])
    assert onum.backward_error(A, X, B) <= _bwd_tol(prec)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ldl", [False, True])
def test_two_unconnected_laplacians(ldl, prec):
    n1, cp1, ri1, v1 = M.laplacian_2d_5pt(24)
    n2, cp2, ri2, v2 = M.diffusion_3d_7pt(8, seed=2)
    rows, cols, vals = [], [], []
    for (n0, cp, ri, v), off in (((n1, cp1, ri1, v1), 0), ((n2, cp2, ri2, v2), n1)):
        c = np.repeat(np.arange(n0), np.diff(cp))
        rows.append(ri.astype(np.int64) + off)
        cols.append(c + off)
        vals.append(v)
    n, cp, ri, v = _lower_csc(n1 + n2, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))
    A, M_, B, X = _solve(n, cp, ri, v, "AMD", ldl, prec=prec)
    S = onum.Structure.of(M_)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    Xo = onum.solve(S, F, B, ldl)
    assert np.linalg.norm(X - Xo) <= _fwd_tol(prec, False) * np.linalg.norm(Xo)
    assert onum.backward_error(A, X, B) <= _bwd_tol(prec)
