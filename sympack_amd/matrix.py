"""Python mirror of symPACK's symPACKMatrix2D interface on one GPU.

Method names, argument meaning, call order and error behaviour follow
include/sympack/symPACKMatrix2D.hpp:2469-2631 and the driver
driver/run_sympack2D.cpp:74-128:

    opts = symPACKOptions(ordering="MMD")
    A = DistSparseMatrix(n, colptr, rowind, nzval)   # lower triangle, CSC
    M = symPACKMatrix2D()
    M.Init(opts)
    M.SymbolicFactorization(A)
    M.DistributeMatrix(A)
    M.Factorize()
    M.Solve(rhs, nrhs)
    x = M.GetSolution(nrhs)

Every numeric call goes through libsympack_amd.so; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class RelaxationParameters:
    """include/sympack/Types.hpp:22-95 (defaults nrelax 8/32/64, zrelax 0.8/0.1/0.05)."""

    def __init__(self, nrelax0=8, nrelax1=32, nrelax2=64, maxSize=0,
                 zrelax0=0.8, zrelax1=0.1, zrelax2=0.05):
        self.nrelax0, self.nrelax1, self.nrelax2 = nrelax0, nrelax1, nrelax2
        self.maxSize = maxSize
        self.zrelax0, self.zrelax1, self.zrelax2 = zrelax0, zrelax1, zrelax2

    def SetMaxSize(self, m):
        self.maxSize = m
        if m > 0:
            self.nrelax0 = min(self.nrelax0, m)
            self.nrelax1 = min(self.nrelax1, m)
            self.nrelax2 = min(self.nrelax2, m)


def refine_code(name):
    """symPACKOptions.order_refinement_str -> spk_options.refine
    (symPACKMatrix2D.hpp:3090-3263): SET = SET31, SET10 .. SET31 = 10*ordflag
    + altflag, TSPB = 40 (tour with skeleton compression), other TSPB* names
    = 41 (without it), anything else = 0 (ordering left as it is)."""
    if name == "TSPB":
        return 40
    if name.startswith("TSPB"):
        return 41
    return {"SET": 31, "SET10": 10, "SET11": 11, "SET20": 20, "SET21": 21,
            "SET30": 30, "SET31": 31}.get(name, 0)


def refine_name(code):
    """spk_options.refine -> an order_refinement_str with that meaning."""
    if code == 40:
        return "TSPB"
    if code == 41:
        return "TSPB1"
    return "SET%d" % code if code else "NO"


class symPACKOptions:
    """The symPACKOptions fields run_sympack2D sets (driver/utils.hpp:227-452),
    with the driver's defaults: MMD, LL, relax (8,32,64), unlimited width."""

    def __init__(self, ordering="MMD", decomposition="LL", relax=None, perm=None,
                 precision="FP64", ir_max_iter=30, ir_tol=1e-12, verbose=0,
                 profile=False, device=0, merge_updates=1, solve_dataflow=False,
                 order_refinement_str="NO", solve_inverse=True, autotune=False, schur_kernel=0):
        self.orderingStr = ordering
        self.decomposition = decomposition
        self.relax = relax if relax is not None else RelaxationParameters()
        self.perm = perm
        self.precision = precision
        self.ir_max_iter = ir_max_iter
        self.ir_tol = ir_tol
        self.verbose = verbose
        self.profile = profile
        self.device = device
        self.merge_updates = merge_updates
        self.solve_dataflow = solve_dataflow
        # supernode reordering (symPACKMatrix2D.hpp:3090-3263): NO, SET,
        # SET10, SET11, SET20, SET21, SET30, SET31, TSPB (tour with skeleton
        # compression), other TSPB* names (tour without it)
        self.order_refinement_str = order_refinement_str
        # partitioned-inverse solve where every supernode is <= 160 wide
        self.solve_inverse = solve_inverse
        # time the Schur launch configurations per etree level on the first
        # factorization (structures with >= 1e9 Schur flops)
        self.autotune = autotune
        # Schur-update kernel on every level: 0 autotuner / plan, 1 64 x 64 tiles staged
        # through LDS, 2 128 x 64 tiles, 3 64 x 64 tiles with one wave per tile
        self.schur_kernel = schur_kernel

    @classmethod
    def from_args(cls, args):
        """Parse run_sympack2D flags (-ordering, -fact, -relax, -b, -precision ...)."""
        lib = _lib.load()
        o = _lib.SpkOptions()
        lib.spk_default_options(ctypes.byref(o))
        arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        _lib.check(lib.spk_options_from_args(ctypes.byref(o), len(args), arr), "options")
        inv = {v: k for k, v in _lib.ORDERINGS.items()}
        r = RelaxationParameters(o.nrelax0, o.nrelax1, o.nrelax2, o.maxsize,
                                 o.zrelax0, o.zrelax1, o.zrelax2)
        return cls(ordering=inv[o.ordering], decomposition="LDL" if o.decomposition else "LL",
                   relax=r, precision="FP32_IR" if o.precision else "FP64",
                   ir_max_iter=o.ir_max_iter, ir_tol=o.ir_tol, verbose=o.verbose,
                   merge_updates=int(o.merge_updates), solve_dataflow=bool(o.solve_dataflow),
                   order_refinement_str=refine_name(o.refine),
                   solve_inverse=bool(o.solve_inverse), autotune=bool(o.autotune),
                   schur_kernel=int(o.schur_kernel))

    def to_c(self):
        lib = _lib.load()
        o = _lib.SpkOptions()
        lib.spk_default_options(ctypes.byref(o))
        if self.orderingStr not in _lib.ORDERINGS:
            raise ValueError(
                "This ordering method is not supported by symPACK. Valid options are:"
                "NATURAL MMD AMD RCM NDBOX NDGRID USER ")
        o.ordering = _lib.ORDERINGS[self.orderingStr]
        o.decomposition = _lib.FACT[self.decomposition]
        o.nrelax0, o.nrelax1, o.nrelax2 = self.relax.nrelax0, self.relax.nrelax1, self.relax.nrelax2
        o.maxsize = self.relax.maxSize
        o.zrelax0, o.zrelax1, o.zrelax2 = self.relax.zrelax0, self.relax.zrelax1, self.relax.zrelax2
        o.precision = _lib.PREC[self.precision]
        o.ir_max_iter = self.ir_max_iter
        o.ir_tol = self.ir_tol
        o.verbose = self.verbose
        o.profile = 1 if self.profile else 0
        o.device = self.device
        o.merge_updates = int(self.merge_updates)
        o.solve_dataflow = 1 if self.solve_dataflow else 0
        o.solve_inverse = 1 if self.solve_inverse else 0
        o.autotune = 1 if self.autotune else 0
        o.schur_kernel = int(self.schur_kernel)
        o.refine = refine_code(self.order_refinement_str)
        return o


class DistSparseMatrix:
    """Lower triangle (diagonal included) in CSC form, as DistSparseMatrix<T>
    holds it on one rank (include/sympack/DistSparseMatrix.hpp:46-61)."""

    def __init__(self, n, colptr, rowind, nzval=None, baseval=None):
        self.size = int(n)
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        self.rowind = np.ascontiguousarray(rowind, dtype=np.int32)
        self.nzval = None if nzval is None else np.ascontiguousarray(nzval, dtype=np.float64)
        self.baseval = int(self.colptr[0]) if baseval is None else baseval
        self.nnz = int(self.colptr[-1] - self.baseval)

    @classmethod
    def from_rb(cls, filename):
        """ReadMatrix(..., "RB"/"HB") -> ReadHB_PARA (utility.hpp:1274-1482)."""
        lib = _lib.load()
        n = ctypes.c_int32()
        nnz = ctypes.c_int64()
        _lib.check(lib.spk_read_rb(filename.encode(), ctypes.byref(n), ctypes.byref(nnz)), "ReadMatrix")
        colptr = np.zeros(n.value + 1, np.int64)
        rowind = np.zeros(nnz.value, np.int32)
        nzval = np.zeros(nnz.value, np.float64)
        _lib.check(lib.spk_read_rb_data(filename.encode(), _ptr(colptr, ctypes.c_int64),
                                        _ptr(rowind, ctypes.c_int32), _ptr(nzval, ctypes.c_double)),
                   "ReadMatrix")
        return cls(n.value, colptr, rowind, nzval, baseval=1)

    @classmethod
    def read(cls, filename, informat="HARWELL_BOEING"):
        """ReadMatrix (utility.hpp:1693-1716): informat "CSC" (binary),
        "HB" / "RB" / "HARWELL_BOEING" or "matrix" (text); other names raise
        ValueError("Unknown matrix format.")."""
        lib = _lib.load()
        n = ctypes.c_int32()
        nnz = ctypes.c_int64()
        f, fmt = filename.encode(), informat.encode()
        _lib.check(lib.spk_read_matrix(f, fmt, ctypes.byref(n), ctypes.byref(nnz)), "ReadMatrix")
        colptr = np.zeros(n.value + 1, np.int64)
        rowind = np.zeros(nnz.value, np.int32)
        nzval = np.zeros(nnz.value, np.float64)
        _lib.check(lib.spk_read_matrix_data(f, fmt, _ptr(colptr, ctypes.c_int64), _ptr(rowind, ctypes.c_int32),
                                            _ptr(nzval, ctypes.c_double)), "ReadMatrix")
        return cls(n.value, colptr, rowind, nzval, baseval=1)

    def write(self, filename, outformat):
        """Write the lower triangle as "CSC", "HB"/"RB"/"HARWELL_BOEING" or "matrix"."""
        lib = _lib.load()
        cp = np.ascontiguousarray(self.colptr, np.int64)
        ri = np.ascontiguousarray(self.rowind, np.int32)
        _lib.check(lib.spk_write_matrix(filename.encode(), outformat.encode(), self.size, _ptr(cp, ctypes.c_int64),
                                        _ptr(ri, ctypes.c_int32),
                                        _ptr(np.ascontiguousarray(self.nzval), ctypes.c_double)), "write")

    def write_rb(self, filename, title="sympack_amd"):
        lib = _lib.load()
        _lib.check(lib.spk_write_rb(filename.encode(), self.size, _ptr(self.colptr, ctypes.c_int64),
                                    _ptr(self.rowind, ctypes.c_int32),
                                    _ptr(self.nzval, ctypes.c_double), title.encode()), "write_rb")

    def to_scipy_full(self):
        import scipy.sparse as sp
        b = self.baseval
        low = sp.csc_matrix((self.nzval, self.rowind - b, self.colptr - b), shape=(self.size, self.size))
        return (low + sp.tril(low, -1).T).tocsr()

    def matvec(self, x):
        """Symmetric product from the lower triangle (driver/utils.hpp:95-142)."""
        b = self.baseval
        x = np.asarray(x, dtype=np.float64)
        two_d = x.ndim == 2
        X = x if two_d else x[:, None]
        Y = np.zeros_like(X)
        cols = np.repeat(np.arange(self.size), np.diff(self.colptr))
        rows = self.rowind.astype(np.int64) - b
        v = self.nzval
        np.add.at(Y, rows, v[:, None] * X[cols])
        off = rows != cols
        np.add.at(Y, cols[off], v[off, None] * X[rows[off]])
        return Y if two_d else Y[:, 0]


class symPACKMatrix2D:
    """symPACKMatrix2D<Ptr,Idx,double> on one GPU (symPACKMatrix2D.hpp:2469-2631)."""

    def __init__(self):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.spk_create(ctypes.byref(h)), "create")
        self._h = h
        self._n = 0
        self._nrhs = 0
        self._perm_keepalive = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.spk_destroy(h)
            self._h = None

    # -- symPACKMatrix2D API ------------------------------------------------
    def Init(self, options):
        self.options_ = options
        o = options.to_c()
        if options.orderingStr == "USER":
            if options.perm is None:
                raise ValueError("When using USER, symPACKOptions.perm must be provided.\n")
            self._perm_keepalive = np.ascontiguousarray(options.perm, dtype=np.int32)
            o.perm = _ptr(self._perm_keepalive, ctypes.c_int)
        self._opts_c = o
        _lib.check(self._lib.spk_init(self._h, ctypes.byref(o)), "Init")

    def SymbolicFactorization(self, A):
        self._n = A.size
        _lib.check(self._lib.spk_symbolic_factorization(
            self._h, A.size, _ptr(A.colptr, ctypes.c_int64), _ptr(A.rowind, ctypes.c_int32),
            A.baseval), "SymbolicFactorization")

    def DistributeMatrix(self, A):
        _lib.check(self._lib.spk_distribute_matrix(self._h, _ptr(A.nzval, ctypes.c_double)),
                   "DistributeMatrix")

    def DistributeMatrixDevice(self, dptr):
        """Values already in HBM (device pointer as int)."""
        _lib.check(self._lib.spk_distribute_matrix_device(self._h, ctypes.c_void_p(dptr)),
                   "DistributeMatrix")

    def Factorize(self):
        rc = self._lib.spk_factorize(self._h)
        if rc > 0:
            raise _lib.NotHPDError(rc)
        _lib.check(rc, "Factorize")

    def Solve(self, RHS, nrhs=None, rhs_size=None):
        rhs = np.asarray(RHS, dtype=np.float64)
        if nrhs is None:
            nrhs = 1 if rhs.ndim == 1 else rhs.shape[1]
        # column-major n x nrhs, as the reference's RHS buffer
        flat = np.ascontiguousarray(rhs.ravel(order="F") if rhs.ndim == 2 else rhs)
        if flat.size != self._n * nrhs:
            raise ValueError("RHS has %d entries, expected n*nrhs = %d" % (flat.size, self._n * nrhs))
        _lib.check(self._lib.spk_solve(self._h, _ptr(flat, ctypes.c_double), int(nrhs),
                                       int(flat.size)), "Solve")
        self._nrhs = int(nrhs)

    def SolveDevice(self, dptr, nrhs):
        _lib.check(self._lib.spk_solve_device(self._h, ctypes.c_void_p(dptr), int(nrhs)), "Solve")
        self._nrhs = int(nrhs)

    def SolveDeviceNoAlloc(self, d_rhs, d_x, nrhs):
        """Solve + GetSolution of an FP64 factor from device memory into
        device memory without any device allocation (spk_solve_device_noalloc:
        work vectors in the library's static 128 MiB per-device scratch)."""
        _lib.check(self._lib.spk_solve_device_noalloc(self._h, ctypes.c_void_p(d_rhs), ctypes.c_void_p(d_x),
                                                      int(nrhs)), "Solve")

    def GetSolution(self, nrhs=None):
        nrhs = self._nrhs if nrhs is None else nrhs
        out = np.zeros(self._n * nrhs, np.float64)
        _lib.check(self._lib.spk_get_solution(self._h, _ptr(out, ctypes.c_double), int(nrhs)),
                   "GetSolution")
        return out.reshape(nrhs, self._n).T.copy() if nrhs > 1 else out

    def GetSolutionDevice(self, dptr, nrhs):
        _lib.check(self._lib.spk_get_solution_device(self._h, ctypes.c_void_p(dptr), int(nrhs)),
                   "GetSolution")

    # -- accessors (symPACKMatrixBase.hpp:56-103) ---------------------------
    def stats(self):
        st = _lib.SpkStats()
        _lib.check(self._lib.spk_get_stats(self._h, ctypes.byref(st)), "stats")
        return {f: getattr(st, f) for f, _ in _lib.SpkStats._fields_}

    def kernel_stats(self, kind):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        _lib.check(self._lib.spk_get_kernel_stats(self._h, kind, ctypes.byref(n), ctypes.byref(ms)),
                   "kernel_stats")
        return n.value, ms.value

    def reset_kernel_stats(self):
        self._lib.spk_reset_kernel_stats(self._h)

    def launch_trace(self, max_records=100000):
        k = np.zeros(max_records, np.int32)
        g = np.zeros(max_records, np.int64)
        ms = np.zeros(max_records, np.float64)
        cnt = _lib.check(self._lib.spk_get_launch_trace(self._h, _ptr(k, ctypes.c_int32),
                                                        _ptr(g, ctypes.c_int64),
                                                        _ptr(ms, ctypes.c_double), max_records), "trace")
        return k[:cnt], g[:cnt], ms[:cnt]

    def autotune_log(self, max_records=4096):
        """The autotuner's record per Schur launch: dicts with level, split,
        order (bit 0: longest first instead of plan order, bit 1: 128 x 64
        output tiles, bit 2: 64 x 64 tiles with one wave per tile), ms_default
        and ms_best."""
        lv, sp_, od = (np.zeros(max_records, np.int32) for _ in range(3))
        m0, m1 = np.zeros(max_records), np.zeros(max_records)
        cnt = _lib.check(self._lib.spk_get_autotune(self._h, _ptr(lv, ctypes.c_int32), _ptr(sp_, ctypes.c_int32),
                                                   _ptr(od, ctypes.c_int32), _ptr(m0, ctypes.c_double),
                                                   _ptr(m1, ctypes.c_double), max_records), "autotune")
        return [{"level": int(lv[i]), "split": int(sp_[i]), "order": int(od[i]), "ms_default": float(m0[i]),
                 "ms_best": float(m1[i])} for i in range(cnt)]

    def launch_levels(self, max_records=100000):
        """Etree level of every record of launch_trace() (-1 outside the level sequences)."""
        lv = np.zeros(max_records, np.int32)
        cnt = _lib.check(self._lib.spk_get_launch_levels(self._h, _ptr(lv, ctypes.c_int32), max_records),
                         "launch levels")
        return lv[:cnt]

    def supernode_levels(self):
        """Etree level of every supernode (0 = leaves), in supernode order."""
        lv = np.zeros(self.TotalSupernodeCnt(), np.int32)
        _lib.check(self._lib.spk_get_supernode_levels(self._h, _ptr(lv, ctypes.c_int32)), "supernode levels")
        return lv

    def set_profile(self, on, kinds=None):
        """Time kernel launches with HIP events (all kinds, or the given kinds)."""
        if kinds is None:
            _lib.check(self._lib.spk_set_profile(self._h, 1 if on else 0), "set_profile")
        else:
            mask = 0
            for k in kinds:
                mask |= 1 << k
            _lib.check(self._lib.spk_set_profile_mask(self._h, 1 if on else 0, mask), "set_profile")

    def DeviceBytes(self, nrhs=1):
        """HBM bytes held after DistributeMatrix plus solve buffers for nrhs."""
        b = ctypes.c_int64()
        _lib.check(self._lib.spk_device_bytes(self._h, int(nrhs), ctypes.byref(b)), "DeviceBytes")
        return b.value

    def ReleaseDevice(self):
        """Free the device memory; DistributeMatrix must be called again."""
        _lib.check(self._lib.spk_release_device(self._h), "ReleaseDevice")

    def set_graphs(self, on):
        """Replay launch sequences from captured HIP graphs (default) or not."""
        _lib.check(self._lib.spk_set_graphs(self._h, 1 if on else 0), "set_graphs")

    def TotalSupernodeCnt(self):
        return self.stats()["nsuper"]

    def GetSupernodalPartition(self):
        xs = np.zeros(self.TotalSupernodeCnt() + 1, np.int32)
        _lib.check(self._lib.spk_get_supernodal_partition(self._h, _ptr(xs, ctypes.c_int32)), "xsuper")
        return xs

    def GetSupMembership(self):
        s = np.zeros(self._n, np.int32)
        _lib.check(self._lib.spk_get_sup_membership(self._h, _ptr(s, ctypes.c_int32)), "supmem")
        return s

    def GetOrdering(self):
        perm = np.zeros(self._n, np.int32)
        invp = np.zeros(self._n, np.int32)
        _lib.check(self._lib.spk_get_ordering(self._h, _ptr(perm, ctypes.c_int32),
                                              _ptr(invp, ctypes.c_int32)), "ordering")
        return perm, invp

    def GetETree(self):
        p = np.zeros(self._n, np.int32)
        _lib.check(self._lib.spk_get_etree(self._h, _ptr(p, ctypes.c_int32)), "etree")
        return p

    def GetColumnCounts(self):
        a = np.zeros(self._n, np.int32)
        b = np.zeros(self._n, np.int32)
        _lib.check(self._lib.spk_get_column_counts(self._h, _ptr(a, ctypes.c_int32),
                                                   _ptr(b, ctypes.c_int32)), "cc")
        return a, b

    def GetStructure(self):
        nnz = ctypes.c_int64()
        _lib.check(self._lib.spk_get_structure_size(self._h, ctypes.byref(nnz)), "lindx")
        xl = np.zeros(self.TotalSupernodeCnt() + 1, np.int64)
        lx = np.zeros(nnz.value, np.int32)
        _lib.check(self._lib.spk_get_structure(self._h, _ptr(xl, ctypes.c_int64),
                                               _ptr(lx, ctypes.c_int32)), "lindx")
        return xl, lx

    def GetUpdateMerges(self):
        """deferred_into[I] (0-based): the parent that computes supernode I's
        update outside the parent's columns, -1 where I scatters all of it."""
        d = np.zeros(self.TotalSupernodeCnt(), np.int32)
        _lib.check(self._lib.spk_get_update_merges(self._h, _ptr(d, ctypes.c_int32)), "merges")
        return d

    def ValidateSolvePlan(self):
        """Raise if the dependency-driven solve's task lists are inconsistent."""
        _lib.check(self._lib.spk_validate_solve_plan(self._h), "solve plan")

    def GetCells(self):
        nc = ctypes.c_int64()
        nb = ctypes.c_int64()
        _lib.check(self._lib.spk_get_cell_counts(self._h, ctypes.byref(nc), ctypes.byref(nb)), "cells")
        a = [np.zeros(nc.value, np.int32) for _ in range(4)]
        b = [np.zeros(nb.value, np.int32) for _ in range(2)]
        _lib.check(self._lib.spk_get_cells(self._h, *[_ptr(x, ctypes.c_int32) for x in a + b]), "cells")
        return dict(K=a[0], I=a[1], nrows=a[2], nblocks=a[3], block_first_row=b[0], block_nrows=b[1])

    def GetFactorPanel(self, I):
        """Panel of supernode I (1-based): structure rows x columns."""
        r = ctypes.c_int32()
        w = ctypes.c_int32()
        _lib.check(self._lib.spk_get_panel_dims(self._h, I, ctypes.byref(r), ctypes.byref(w)), "panel")
        out = np.zeros(r.value * w.value, np.float64)
        _lib.check(self._lib.spk_get_factor_panel(self._h, I, _ptr(out, ctypes.c_double)), "panel")
        return out.reshape(w.value, r.value).T.copy()


def compute_ordering(ordering, A):
    """The ordering routine alone (symPACK::Ordering, src/Ordering.cpp:103-525):
    returns (perm, invp), 1-based, before the elimination-tree postorder."""
    lib = _lib.load()
    perm = np.zeros(A.size, np.int32)
    invp = np.zeros(A.size, np.int32)
    _lib.check(lib.spk_compute_ordering(_lib.ORDERINGS[ordering], A.size, _ptr(A.colptr, ctypes.c_int64),
                                        _ptr(A.rowind, ctypes.c_int32), A.baseval,
                                        _ptr(perm, ctypes.c_int32), _ptr(invp, ctypes.c_int32)),
               "ordering")
    return perm, invp
