"""Oracle: ctypes front-end of liboracle.so (oracle/supernodal.c), the CPU
restatement of symPACK's supernodal numeric factorization and solve.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker, never by the product path.

The structure (perm, xsuper, xlindx, lindx) is taken from the symbolic
analysis under test; that analysis is pinned separately against the
reference's Fortran orderings and the Python restatement in
oracle/symbolic.py.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run make -C oracle")
        L = ctypes.CDLL(path)
        P32 = ctypes.POINTER(ctypes.c_int32)
        P64 = ctypes.POINTER(ctypes.c_int64)
        PD = ctypes.POINTER(ctypes.c_double)
        L.oracle_factor_size.restype = ctypes.c_int64
        L.oracle_factor_size.argtypes = [ctypes.c_int, ctypes.c_int, P32, P64]
        L.oracle_factor.restype = ctypes.c_int
        L.oracle_factor.argtypes = [ctypes.c_int, ctypes.c_int, P32, P64, P32, P32, P64, P32, PD,
                                    ctypes.c_int, PD]
        L.oracle_solve.restype = ctypes.c_int
        L.oracle_solve.argtypes = [ctypes.c_int, ctypes.c_int, P32, P64, P32, P32, ctypes.c_int, PD,
                                   PD, ctypes.c_int, PD]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Structure:
    def __init__(self, perm, xsuper, xlindx, lindx):
        self.perm = np.ascontiguousarray(perm, np.int32)
        self.xsuper = np.ascontiguousarray(xsuper, np.int32)
        self.xlindx = np.ascontiguousarray(xlindx, np.int64)
        self.lindx = np.ascontiguousarray(lindx, np.int32)
        self.n = len(self.perm)
        self.nsuper = len(self.xsuper) - 1

    @classmethod
    def of(cls, M):
        perm, _ = M.GetOrdering()
        xl, lx = M.GetStructure()
        return cls(perm, M.GetSupernodalPartition(), xl, lx)


def factor(S, A, ldl=False):
    """Returns (info, factor storage).  A: DistSparseMatrix (lower CSC)."""
    L = lib()
    size = L.oracle_factor_size(S.n, S.nsuper, _p(S.xsuper, ctypes.c_int32), _p(S.xlindx, ctypes.c_int64))
    F = np.zeros(size, np.float64)
    cp = np.ascontiguousarray(A.colptr - A.baseval, np.int64)
    ri = np.ascontiguousarray(A.rowind - A.baseval, np.int32)
    info = L.oracle_factor(S.n, S.nsuper, _p(S.xsuper, ctypes.c_int32), _p(S.xlindx, ctypes.c_int64),
                           _p(S.lindx, ctypes.c_int32), _p(S.perm, ctypes.c_int32),
                           _p(cp, ctypes.c_int64), _p(ri, ctypes.c_int32),
                           _p(np.ascontiguousarray(A.nzval), ctypes.c_double), 1 if ldl else 0,
                           _p(F, ctypes.c_double))
    return info, F


def panel(S, F, I):
    """Supernode I (1-based) as an (r x w) array."""
    off = 0
    for J in range(I - 1):
        off += (S.xlindx[J + 1] - S.xlindx[J]) * (S.xsuper[J + 1] - S.xsuper[J])
    r = int(S.xlindx[I] - S.xlindx[I - 1])
    w = int(S.xsuper[I] - S.xsuper[I - 1])
    return F[off:off + r * w].reshape(r, w)


def panels(S, F):
    out = []
    off = 0
    for I in range(S.nsuper):
        r = int(S.xlindx[I + 1] - S.xlindx[I])
        w = int(S.xsuper[I + 1] - S.xsuper[I])
        out.append(F[off:off + r * w].reshape(r, w))
        off += r * w
    return out


def solve(S, F, B, ldl=False):
    B = np.asarray(B, np.float64)
    two = B.ndim == 2
    Bm = B if two else B[:, None]
    nrhs = Bm.shape[1]
    b = np.ascontiguousarray(Bm.ravel(order="F"))
    x = np.zeros_like(b)
    lib().oracle_solve(S.n, S.nsuper, _p(S.xsuper, ctypes.c_int32), _p(S.xlindx, ctypes.c_int64),
                       _p(S.lindx, ctypes.c_int32), _p(S.perm, ctypes.c_int32), 1 if ldl else 0,
                       _p(F, ctypes.c_double), _p(b, ctypes.c_double), nrhs, _p(x, ctypes.c_double))
    X = x.reshape(nrhs, S.n).T
    return X if two else X[:, 0]


def factor_and_solve(M, A, B, ldl=False):
    S = Structure.of(M)
    info, F = factor(S, A, ldl)
    if info != 0:
        raise RuntimeError("oracle factorization failed, info=%d" % info)
    return solve(S, F, B, ldl)


def backward_error(A, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)."""
    X = np.asarray(X).reshape(A.size, -1)
    B = np.asarray(B).reshape(A.size, -1)
    R = B - A.matvec(X)
    absA = A.__class__(A.size, A.colptr, A.rowind, np.abs(A.nzval), A.baseval)
    normA = np.max(absA.matvec(np.ones(A.size)))
    num = np.max(np.abs(R), axis=0)
    den = normA * np.max(np.abs(X), axis=0) + np.max(np.abs(B), axis=0)
    return float(np.max(num / den))


def _blas_pointers():
    """Addresses of dpotrf, dtrsm, dsyrk in scipy's LAPACK/BLAS (the
    Fortran entry points its cython_blas / cython_lapack modules export)."""
    import scipy.linalg.cython_blas as cb
    import scipy.linalg.cython_lapack as cl
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    name = ctypes.pythonapi.PyCapsule_GetName
    name.restype = ctypes.c_char_p
    name.argtypes = [ctypes.py_object]

    def ptr(cap):
        return get(cap, name(cap))
    return ptr(cl.__pyx_capi__["dpotrf"]), ptr(cb.__pyx_capi__["dtrsm"]), ptr(cb.__pyx_capi__["dsyrk"])


def factor_blas(S, A, threads=1):
    """BLAS-backed LL^T of the same structure (oracle_factor_blas): returns
    (info, factor storage, seconds of the factorization proper, assembly
    excluded).  threads: BLAS threads (the scatter runs on the calling thread)."""
    import time
    from threadpoolctl import threadpool_limits
    L = lib()
    P32 = ctypes.POINTER(ctypes.c_int32)
    P64 = ctypes.POINTER(ctypes.c_int64)
    PD = ctypes.POINTER(ctypes.c_double)
    L.oracle_assemble.restype = ctypes.c_int
    L.oracle_assemble.argtypes = [ctypes.c_int, ctypes.c_int, P32, P64, P32, P32, P64, P32, PD, PD]
    L.oracle_factor_blas.restype = ctypes.c_int
    L.oracle_factor_blas.argtypes = [ctypes.c_int, ctypes.c_int, P32, P64, P32, PD, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
    size = L.oracle_factor_size(S.n, S.nsuper, _p(S.xsuper, ctypes.c_int32), _p(S.xlindx, ctypes.c_int64))
    F = np.zeros(size, np.float64)
    cp = np.ascontiguousarray(A.colptr - A.baseval, np.int64)
    ri = np.ascontiguousarray(A.rowind - A.baseval, np.int32)
    rc = L.oracle_assemble(S.n, S.nsuper, _p(S.xsuper, ctypes.c_int32), _p(S.xlindx, ctypes.c_int64),
                           _p(S.lindx, ctypes.c_int32), _p(S.perm, ctypes.c_int32), _p(cp, ctypes.c_int64),
                           _p(ri, ctypes.c_int32), _p(np.ascontiguousarray(A.nzval), ctypes.c_double),
                           _p(F, ctypes.c_double))
    if rc != 0:
        raise RuntimeError("assembly failed")
    potrf, trsm, syrk = _blas_pointers()
    with threadpool_limits(limits=threads, user_api="blas"):
        t0 = time.perf_counter()
        info = L.oracle_factor_blas(S.n, S.nsuper, _p(S.xsuper, ctypes.c_int32), _p(S.xlindx, ctypes.c_int64),
                                    _p(S.lindx, ctypes.c_int32), _p(F, ctypes.c_double), potrf, trsm, syrk)
        dt = time.perf_counter() - t0
    return info, F, dt
