"""C ABI surface (every declared symbol is exported and bound), the option
parser, Rutherford-Boeing I/O and the matrix generators (CPU only)."""
import os
import re

import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import _lib
from sympack_amd import matrices as M

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "sympack_amd.h")


def test_every_declared_symbol_is_exported_and_bound():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(spk_\w+|symPACK_\w+)\s*\(", text))
    lib = _lib.load()
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(lib, n), n
        assert n in bound, n


def test_rank_and_init():
    lib = _lib.load()
    import ctypes
    r = ctypes.c_int(-1)
    assert lib.symPACK_Init(None, None) == 0
    assert lib.symPACK_Rank(ctypes.byref(r)) == 0 and r.value == 0
    assert lib.symPACK_Finalize() == 0


def test_options_from_args_like_run_sympack2D():
    o = sp.symPACKOptions.from_args(["-in", "x.rb", "-ordering", "NDBOX", "-fact", "LDL", "-relax", "4", "16",
                                     "32", "-b", "100", "-precision", "FP32"])
    assert o.orderingStr == "NDBOX" and o.decomposition == "LDL"
    assert (o.relax.nrelax0, o.relax.nrelax1, o.relax.nrelax2, o.relax.maxSize) == (4, 16, 32, 100)
    assert o.precision == "FP32_IR"
    o = sp.symPACKOptions.from_args(["-relax", "off"])
    assert o.relax.nrelax0 == 0
    o = sp.symPACKOptions.from_args([])
    assert o.orderingStr == "MMD" and o.relax.maxSize == 0


@pytest.mark.parametrize("flag,want", [("FP32_IR", "FP32_IR"), ("FP32IR", "FP32_IR"), ("fp32", "FP32_IR"),
                                       ("FP64", "FP64")])
def test_precision_flag_spellings(flag, want):
    assert sp.symPACKOptions.from_args(["-precision", flag]).precision == want


def test_precision_flag_rejects_unknown_values():
    with pytest.raises(ValueError, match="precision"):
        sp.symPACKOptions.from_args(["-precision", "FP16"])


def test_refine_code_outside_the_reordering_modes_is_rejected():
    import ctypes
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.spk_create(ctypes.byref(h)) == 0
    try:
        o = _lib.SpkOptions()
        lib.spk_default_options(ctypes.byref(o))
        for good in (0, 10, 11, 20, 21, 30, 31, 40, 41):
            o.refine = good
            assert lib.spk_init(h, ctypes.byref(o)) == 0
        for bad in (12, 42, 50, 1, -5):
            o.refine = bad
            assert lib.spk_init(h, ctypes.byref(o)) == -2
            assert b"refine" in lib.spk_last_error()
    finally:
        lib.spk_destroy(h)


def test_rb_round_trip(tmp_path):
    n, cp, ri, v = M.diffusion_3d_7pt(5, seed=3)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    f = str(tmp_path / "a.rb")
    A.write_rb(f)
    B = sp.DistSparseMatrix.from_rb(f)
    assert B.size == n and B.baseval == 1
    assert np.array_equal(B.colptr - 1, cp)
    assert np.array_equal(B.rowind - 1, ri)
    assert np.array_equal(B.nzval, v)  # 17 significant digits: exact


def test_rb_reads_reference_style_file(tmp_path):
    # 3x3 lower triangle, fixed-width records as ReadHB_PARA expects
    f = tmp_path / "small.rb"
    f.write_text(
        "small test matrix                                                       KEY     \n"
        "             3             1             1             1             0\n"
        "rsa                        3             3             5             0\n"
        "(4I5)           (5I5)           (5E15.6)            \n"
        "    1    3    5    6\n"
        "    3    1    2    3    3\n"
        "   4.000000E+00  -1.000000E+00   4.000000E+00  -1.000000E+00   4.000000E+00\n")
    A = sp.DistSparseMatrix.from_rb(str(f))
    assert A.size == 3
    assert list(A.colptr) == [1, 3, 5, 6]
    assert list(A.rowind) == [1, 3, 2, 3, 3]  # rows sorted per column
    assert list(A.nzval) == [-1.0, 4.0, 4.0, -1.0, 4.0]


def test_generators_sizes_and_symmetry():
    n, cp, ri, v = M.laplacian_2d_5pt(64)
    assert n == 4096 and cp[-1] == 12160
    n, cp, ri, v = M.diffusion_3d_7pt(48)
    assert n == 110592 and cp[-1] == 435456
    n, cp, ri, v = M.shifted_laplacian_3d_7pt(80)
    assert n == 512000 and cp[-1] == 2028800
    n, cp, ri, v = M.stencil_3d_27pt(12)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    F = A.to_scipy_full()
    assert abs(F - F.T).max() == 0
    d = F.diagonal()
    off = np.asarray(abs(F).sum(axis=1)).ravel() - d
    assert np.all(d > off)  # strictly diagonally dominant


def test_cpp_adapter_symbolic_and_errors():
    """include/sympack_amd.hpp compiles against types carrying the
    reference's member names; Init + SymbolicFactorization, the accessors
    and the logic_error paths run without a GPU."""
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "adapter_c1")
    assert os.path.exists(exe), "adapter test not built (make -C sympack_amd/csrc)"
    out = subprocess.run([exe, "symbolic"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "adapter_c1: OK" in out.stdout


@pytest.mark.parametrize("fmt", ["CSC", "matrix", "RB", "HB", "HARWELL_BOEING"])
def test_matrix_formats_round_trip(tmp_path, fmt):
    n, cp, ri, v = M.diffusion_3d_7pt(6, seed=4)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    f = str(tmp_path / ("a." + fmt))
    A.write(f, fmt)
    B = sp.DistSparseMatrix.read(f, fmt)
    assert B.size == n and B.baseval == 1
    assert np.array_equal(B.colptr - 1, cp)
    assert np.array_equal(B.rowind - 1, ri)
    assert np.array_equal(B.nzval, v)


def test_csc_binary_layout_and_row_sorting(tmp_path):
    """The byte layout ParaReadDistSparseMatrix's offsets read
    (utility.hpp:1087-1190): int32 n, nnz, then count + colptr, count + rowind,
    count + nzval; rows come back sorted within each column, values following."""
    f = tmp_path / "m.csc"
    n, nnz = 3, 5
    colptr = np.array([1, 3, 5, 6], np.int32)
    rowind = np.array([3, 1, 3, 2, 3], np.int32)  # column 1 unsorted
    nzval = np.array([-1.0, 4.0, -2.0, 5.0, 6.0])
    with open(f, "wb") as fh:
        np.array([n, nnz, n + 1], np.int32).tofile(fh)
        colptr.tofile(fh)
        np.array([nnz], np.int32).tofile(fh)
        rowind.tofile(fh)
        np.array([nnz], np.int32).tofile(fh)
        nzval.tofile(fh)
    A = sp.DistSparseMatrix.read(str(f), "CSC")
    assert list(A.colptr) == [1, 3, 5, 6]
    assert list(A.rowind) == [1, 3, 2, 3, 3]
    assert list(A.nzval) == [4.0, -1.0, 5.0, -2.0, 6.0]


def test_matrix_text_full_symmetric_is_reduced_to_lower(tmp_path):
    """A "matrix" file with entries above the diagonal is the full symmetric
    matrix: the lower triangle is kept (ToLowerTriangular, impl:310-340)."""
    f = tmp_path / "full.matrix"
    # [[4,-1,0],[-1,4,-2],[0,-2,6]] stored in full
    f.write_text("3 3\n7 0\n1 3 6 8\n1 2\n1 2 3\n2 3\n4 -1\n-1 4 -2\n-2 6\n")
    A = sp.DistSparseMatrix.read(str(f), "matrix")
    assert list(A.colptr) == [1, 3, 5, 6]
    assert list(A.rowind) == [1, 2, 2, 3, 3]
    assert list(A.nzval) == [4.0, -1.0, 4.0, -2.0, 6.0]


def test_unknown_matrix_format():
    with pytest.raises(ValueError, match="Unknown matrix format"):
        sp.DistSparseMatrix.read("/nonexistent", "MATRIX_MARKET")


def test_schur_kernel_option_is_validated():
    m = sp.symPACKMatrix2D()
    for ok in (0, 1, 2, 3):
        m.Init(sp.symPACKOptions(schur_kernel=ok))
    for bad in (4, -1, 64):
        with pytest.raises(ValueError, match="schur_kernel"):
            m.Init(sp.symPACKOptions(schur_kernel=bad))
    o = sp.symPACKOptions.from_args(["-schur_kernel", "2", "-autotune", "1"])
    assert o.schur_kernel == 2 and o.autotune is True
    # tuning costs tens of factorizations: off unless asked for
    assert sp.symPACKOptions().autotune is False
    assert sp.symPACKOptions.from_args([]).autotune is False
