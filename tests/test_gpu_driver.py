"""The run_sympack2D-compatible driver on the 64x64 5-point Laplacian written
as a Rutherford-Boeing file, MMD ordering, nrhs = 1 (the reference's own
CPU case): same command line, relative residual printed at the end."""
import os
import re
import subprocess

import pytest

import sympack_amd as sp
from sympack_amd import matrices as M

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sympack_amd",
                      "run_sympack2D")


@pytest.mark.parametrize("ordering", ["MMD", "AMD", "RCM"])
def test_driver_laplacian_64(tmp_path, ordering):
    assert os.path.exists(DRIVER), "driver not built"
    n, cp, ri, v = M.laplacian_2d_5pt(64)
    f = str(tmp_path / "lap64.rb")
    sp.DistSparseMatrix(n, cp, ri, v).write_rb(f)
    out = subprocess.run([DRIVER, "-in", f, "-inf", "RB", "-ordering", ordering, "-nrhs", "1", "-v", "1"],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    assert "Ordering: %s" % ordering in out.stdout
    res = float(re.search(r"Norm of residual after factorization/solve is (\S+)", out.stdout).group(1))
    assert res < 1e-12
    flops = float(re.search(r"Flops: (\S+)", out.stdout).group(1))
    assert flops > 0


@pytest.mark.parametrize("fmt", ["CSC", "matrix"])
def test_driver_reads_csc_and_matrix_formats(tmp_path, fmt):
    n, cp, ri, v = M.laplacian_2d_5pt(64)
    f = str(tmp_path / ("lap64." + fmt))
    sp.DistSparseMatrix(n, cp, ri, v).write(f, fmt)
    out = subprocess.run([DRIVER, "-in", f, "-inf", fmt, "-ordering", "MMD", "-nrhs", "1"],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    res = float(re.search(r"Norm of residual after factorization/solve is (\S+)", out.stdout).group(1))
    assert res < 1e-12


def test_driver_rejects_unknown_format(tmp_path):
    out = subprocess.run([DRIVER, "-in", str(tmp_path / "x"), "-inf", "MATRIX_MARKET"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode != 0


ADAPTER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "adapter_c1")


def test_cpp_adapter_full_sequence_on_c1():
    """include/sympack_amd.hpp compiled into a C++ program that runs Init ->
    SymbolicFactorization -> DistributeMatrix -> Factorize -> Solve ->
    GetSolution on the 64x64 5-point Laplacian (MMD), LL^T and LDL^T, and
    the not-HPD error."""
    assert os.path.exists(ADAPTER), "adapter test not built"
    out = subprocess.run([ADAPTER, "full"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    assert "adapter_c1: OK" in out.stdout
    assert out.stdout.count("ok   ") >= 11
