"""ctypes binding of libsympack_amd.so (the C ABI in include/sympack_amd.h).

The shared library is built in-tree by ``__graft_entry__.build()``
(``make -C sympack_amd/csrc``).  Importing this module does not touch the GPU;
loading fails loudly if the library has not been built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsympack_amd.so")

ORDERINGS = {"NATURAL": 0, "RCM": 1, "MMD": 2, "AMD": 3, "NDBOX": 4, "NDGRID": 5, "USER": 10}
FACT = {"LL": 0, "LDL": 1}
PREC = {"FP64": 0, "FP32_IR": 1}


class SpkOptions(ctypes.Structure):
    _fields_ = [
        ("ordering", ctypes.c_int),
        ("decomposition", ctypes.c_int),
        ("nrelax0", ctypes.c_int),
        ("nrelax1", ctypes.c_int),
        ("nrelax2", ctypes.c_int),
        ("maxsize", ctypes.c_int),
        ("zrelax0", ctypes.c_double),
        ("zrelax1", ctypes.c_double),
        ("zrelax2", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("ir_max_iter", ctypes.c_int),
        ("ir_tol", ctypes.c_double),
        ("perm", ctypes.POINTER(ctypes.c_int)),
        ("verbose", ctypes.c_int),
        ("profile", ctypes.c_int),
        ("device", ctypes.c_int),
        ("merge_updates", ctypes.c_int),
        ("solve_dataflow", ctypes.c_int),
        ("refine", ctypes.c_int),
        ("solve_inverse", ctypes.c_int),
        ("autotune", ctypes.c_int),
        ("schur_kernel", ctypes.c_int),
    ]


class SpkStats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("nsuper", ctypes.c_int32),
        ("nlevels", ctypes.c_int32),
        ("max_width", ctypes.c_int32),
        ("flops", ctypes.c_double),
        ("nnz_l", ctypes.c_int64),
        ("factor_entries", ctypes.c_int64),
        ("schur_flops", ctypes.c_double),
        ("inpanel_flops", ctypes.c_double),
        ("diag_flops", ctypes.c_double),
        ("schur_entries", ctypes.c_int64),
        ("launches_factor", ctypes.c_int64),
        ("launches_solve", ctypes.c_int64),
        ("factor_ms", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("ir_iterations", ctypes.c_int),
        ("backward_error", ctypes.c_double),
        ("fp32_shift", ctypes.c_double),
        ("ir_converged", ctypes.c_int),
    ]


# (name, restype, argtypes) of every exported entry point
_P = ctypes.c_void_p
_I = ctypes.c_int
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_PD = ctypes.POINTER(ctypes.c_double)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = [
    ("symPACK_Init", _I, [_P, _P]),
    ("symPACK_Finalize", _I, []),
    ("symPACK_Rank", _I, [ctypes.POINTER(ctypes.c_int)]),
    ("symPACK_device_setup", _I, [_I]),
    ("spk_default_options", None, [ctypes.POINTER(SpkOptions)]),
    ("spk_options_from_args", _I, [ctypes.POINTER(SpkOptions), _I, ctypes.POINTER(ctypes.c_char_p)]),
    ("spk_create", _I, [ctypes.POINTER(_P)]),
    ("spk_destroy", _I, [_P]),
    ("spk_init", _I, [_P, ctypes.POINTER(SpkOptions)]),
    ("spk_symbolic_factorization", _I, [_P, _I32, _PI64, _PI32, _I]),
    ("spk_distribute_matrix", _I, [_P, _PD]),
    ("spk_distribute_matrix_device", _I, [_P, _P]),
    ("spk_factorize", _I, [_P]),
    ("spk_solve", _I, [_P, _PD, _I, _I64]),
    ("spk_solve_device", _I, [_P, _P, _I]),
    ("spk_get_solution", _I, [_P, _PD, _I]),
    ("spk_get_solution_device", _I, [_P, _P, _I]),
    ("spk_solve_device_noalloc", _I, [_P, _P, _P, _I]),
    ("spk_get_stats", _I, [_P, ctypes.POINTER(SpkStats)]),
    ("spk_get_kernel_stats", _I, [_P, _I, _PI64, _PD]),
    ("spk_reset_kernel_stats", _I, [_P]),
    ("spk_set_profile", _I, [_P, _I]),
    ("spk_set_profile_mask", _I, [_P, _I, ctypes.c_uint]),
    ("spk_set_graphs", _I, [_P, _I]),
    ("spk_device_bytes", _I, [_P, _I, _PI64]),
    ("spk_release_device", _I, [_P]),
    ("spk_get_launch_trace", _I, [_P, _PI32, _PI64, _PD, _I]),
    ("spk_get_launch_levels", _I, [_P, _PI32, _I]),
    ("spk_get_autotune", _I, [_P, _PI32, _PI32, _PI32, _PD, _PD, _I]),
    ("spk_get_supernode_levels", _I, [_P, _PI32]),
    ("spk_get_supernodal_partition", _I, [_P, _PI32]),
    ("spk_get_sup_membership", _I, [_P, _PI32]),
    ("spk_get_ordering", _I, [_P, _PI32, _PI32]),
    ("spk_get_etree", _I, [_P, _PI32]),
    ("spk_get_column_counts", _I, [_P, _PI32, _PI32]),
    ("spk_get_structure_size", _I, [_P, _PI64]),
    ("spk_get_structure", _I, [_P, _PI64, _PI32]),
    ("spk_get_cell_counts", _I, [_P, _PI64, _PI64]),
    ("spk_get_update_merges", _I, [_P, _PI32]),
    ("spk_validate_solve_plan", _I, [_P]),
    ("spk_get_cells", _I, [_P, _PI32, _PI32, _PI32, _PI32, _PI32, _PI32]),
    ("spk_get_panel_dims", _I, [_P, _I32, _PI32, _PI32]),
    ("spk_get_factor_panel", _I, [_P, _I32, _PD]),
    ("spk_compute_ordering", _I, [_I, _I32, _PI64, _PI32, _I, _PI32, _PI32]),
    ("spk_read_rb", _I, [ctypes.c_char_p, _PI32, _PI64]),
    ("spk_read_rb_data", _I, [ctypes.c_char_p, _PI64, _PI32, _PD]),
    ("spk_write_rb", _I, [ctypes.c_char_p, _I32, _PI64, _PI32, _PD, ctypes.c_char_p]),
    ("spk_read_matrix", _I, [ctypes.c_char_p, ctypes.c_char_p, _PI32, _PI64]),
    ("spk_read_matrix_data", _I, [ctypes.c_char_p, ctypes.c_char_p, _PI64, _PI32, _PD]),
    ("spk_write_matrix", _I, [ctypes.c_char_p, ctypes.c_char_p, _I32, _PI64, _PI32, _PD]),
    ("spk_last_error", ctypes.c_char_p, []),
]

_lib = None


def load():
    """Load the HIP extension; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libsympack_amd.so is not built (expected at %s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SymPACKError(RuntimeError):
    pass


class NotHPDError(SymPACKError):
    """Raised when the factorization meets a non-positive pivot (reference:
    runtime_error("Matrix is not HPD."), src/lapack.cpp:293-306)."""

    def __init__(self, column):
        super().__init__("Matrix is not HPD. (pivot column %d)" % column)
        self.column = column


def check(rc, what):
    if rc < 0:
        msg = load().spk_last_error().decode(errors="replace")
        if msg.startswith("logic_error"):
            raise ValueError("%s: %s" % (what, msg))
        raise SymPACKError("%s: %s" % (what, msg))
    return rc
