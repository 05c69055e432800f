"""The allocation-free solve (spk_solve_device_noalloc) against the
allocating one (spk_solve_device + spk_get_solution_device) through the
C ABI: same factors, same right-hand sides, two calls in a row, the second
with fewer right-hand sides (and, on a second handle, a smaller matrix).

Both paths run the same kernels; the level sweeps add partial sums with
atomics in a varying order, so the two agree to rounding, which the
indefinite static LDL^T amplifies (4.5e-12 relative seen on the 12^3
shifted Laplacian): they are compared at the parity bar, 1e-10 relative,
and both must agree with the oracle's solution to that bar.  Also
covered: a handle that never ran the allocating solve (its solve buffers
were never created), both solve kinds (partitioned inverse for widths
<= 160, block substitution for wider supernodes), and the size limit and
FP32_IR rejection (status -2, nothing launched).
"""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _factored(gen, ordering, bmax, ldl=False, precision="FP64"):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    if bmax is not None:
        rp.SetMaxSize(bmax)
    Mx = sp.symPACKMatrix2D()
    Mx.Init(sp.symPACKOptions(ordering=ordering, decomposition="LDL" if ldl else "LL", relax=rp,
                              precision=precision))
    Mx.SymbolicFactorization(A)
    Mx.DistributeMatrix(A)
    Mx.Factorize()
    return A, Mx


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to("cuda:0")


def _alloc_solve(Mx, b_dev, n, nrhs):
    x = torch.zeros(n * nrhs, dtype=torch.float64, device="cuda:0")
    Mx.SolveDevice(b_dev.data_ptr(), nrhs)
    Mx.GetSolutionDevice(x.data_ptr(), nrhs)
    torch.cuda.synchronize()
    return x.cpu().numpy().reshape(nrhs, n).T


def _noalloc_solve(Mx, b_dev, n, nrhs):
    x = torch.full((n * nrhs,), np.nan, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    Mx.SolveDeviceNoAlloc(b_dev.data_ptr(), x.data_ptr(), nrhs)
    torch.cuda.synchronize()
    return x.cpu().numpy().reshape(nrhs, n).T


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [
    # widths <= 150: partitioned-inverse sweeps
    ("diff3d_16_ndbox_b150", lambda: M.diffusion_3d_7pt(16, seed=4), "NDBOX", 150, False),
    # unlimited widths (supernodes > 160 columns): block substitution
    ("lap2d_64_mmd", lambda: M.laplacian_2d_5pt(64), "MMD", None, False),
    # FP64 static LDL^T (block substitution, D scaling)
    ("shifted_lap_12_ldl", lambda: M.shifted_laplacian_3d_7pt(12, 0.3), "NDBOX", 150, True),
]


@pytest.mark.parametrize("name,gen,ordering,bmax,ldl", CASES, ids=[c[0] for c in CASES])
def test_noalloc_solve_matches_allocating_solve(name, gen, ordering, bmax, ldl):
    A, Mx = _factored(gen, ordering, bmax, ldl)
    n = A.size
    S = onum.Structure.of(Mx)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    rng = np.random.default_rng(11)
    for nrhs in (7, 2):  # second call smaller
        print("case %s nrhs=%d" % (name, nrhs), flush=True)
        B = rng.uniform(-1, 1, (n, nrhs))
        b = _dev(B)
        Xn = _noalloc_solve(Mx, b, n, nrhs)  # first: before any allocating solve on this handle
        Xa = _alloc_solve(Mx, b, n, nrhs)
        assert np.all(np.isfinite(Xn))
        assert _rel(Xn, Xa) <= 1e-10, (name, nrhs, _rel(Xn, Xa))
        Xo = onum.solve(S, F, B, ldl)
        assert _rel(Xn, Xo) <= 1e-10, (name, nrhs, _rel(Xn, Xo))


def test_noalloc_smaller_matrix_after_larger():
    """Second call on a smaller matrix: the shared scratch keeps the larger
    call's contents, which must not leak into the result."""
    A1, M1 = _factored(lambda: M.diffusion_3d_7pt(20, seed=1), "NDBOX", 150)
    A2, M2 = _factored(lambda: M.diffusion_3d_7pt(10, seed=2), "NDBOX", 150)
    rng = np.random.default_rng(5)
    for A, Mx, nrhs in ((A1, M1, 9), (A2, M2, 3)):
        n = A.size
        B = rng.uniform(-1, 1, (n, nrhs))
        b = _dev(B)
        Xn = _noalloc_solve(Mx, b, n, nrhs)
        Xa = _alloc_solve(Mx, b, n, nrhs)
        assert _rel(Xn, Xa) <= 1e-10, (n, nrhs, _rel(Xn, Xa))
        assert onum.backward_error(A, Xn, B) <= 1e-13


def test_noalloc_rejects_oversize_and_fp32():
    A, Mx = _factored(lambda: M.diffusion_3d_7pt(8, seed=3), "NDBOX", 150)
    n = A.size
    big = (128 << 20) // (3 * 8 * n) + 1  # 3 n nrhs doubles just above 128 MiB
    b = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError):
        Mx.SolveDeviceNoAlloc(b.data_ptr(), b.data_ptr(), big)
    A32, M32 = _factored(lambda: M.diffusion_3d_7pt(8, seed=3), "NDBOX", 150, precision="FP32_IR")
    bb = torch.zeros(A32.size, dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError):
        M32.SolveDeviceNoAlloc(bb.data_ptr(), bb.data_ptr(), 1)
