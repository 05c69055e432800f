"""Synthetic SPD / symmetric test matrices of the benchmark workloads.

All generators return (n, colptr, rowind, nzval) of the LOWER triangle in
0-based CSC with rows sorted inside every column, the layout
DistSparseMatrix<T> holds (include/sympack/DistSparseMatrix.hpp:46-61).
Grid node (x, y, z) has the natural index x + nx*y + nx*ny*z (x fastest), the
numbering the geometric nested-dissection routines boxnd.f / gridnd.f assume.
Random draws use numpy.random.default_rng(seed) in the order documented in
each function.
"""
import numpy as np


def _lower_csc(n, rows, cols, vals):
    """COO (lower entries only, row >= col) -> sorted 0-based CSC."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64)
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    colptr = np.zeros(n + 1, np.int64)
    np.add.at(colptr, cols + 1, 1)
    colptr = np.cumsum(colptr)
    return n, colptr, rows.astype(np.int32), vals


def _grid_edges(nx, ny, nz, offsets):
    """Undirected grid edges for the given forward offsets; returns
    (lo, hi, offset_id) arrays with lo < hi, offsets enumerated in order."""
    idx = np.arange(nx * ny * nz, dtype=np.int64).reshape(nz, ny, nx)
    los, his, oid = [], [], []
    for k, (dx, dy, dz) in enumerate(offsets):
        xs = slice(max(0, -dx), nx - max(0, dx))
        ys = slice(max(0, -dy), ny - max(0, dy))
        zs = slice(max(0, -dz), nz - max(0, dz))
        a = idx[zs, ys, xs]
        b = idx[zs.start + dz:zs.stop + dz, ys.start + dy:ys.stop + dy, xs.start + dx:xs.stop + dx]
        a = a.ravel()
        b = b.ravel()
        los.append(np.minimum(a, b))
        his.append(np.maximum(a, b))
        oid.append(np.full(a.size, k))
    return np.concatenate(los), np.concatenate(his), np.concatenate(oid)


def laplacian_2d_5pt(k=64):
    """2D 5-point Laplacian on a k x k grid with Dirichlet boundary:
    a_ii = 4, a_ij = -1 for N/S/E/W neighbours (k=64: n=4096, 12160 stored)."""
    n = k * k
    lo, hi, _ = _grid_edges(k, k, 1, [(1, 0, 0), (0, 1, 0)])
    rows = np.concatenate([np.arange(n), hi])
    cols = np.concatenate([np.arange(n), lo])
    vals = np.concatenate([np.full(n, 4.0), np.full(lo.size, -1.0)])
    return _lower_csc(n, rows, cols, vals)


def diffusion_3d_7pt(k=48, seed=0, lo_exp=-2.0, hi_exp=2.0, shift=1e-3):
    """3D 7-point variable-coefficient diffusion on a k^3 grid: one face
    coefficient c = 10**U(lo_exp, hi_exp) per undirected face (drawn in the
    edge order x-faces, y-faces, z-faces), a_ij = -c, a_ii = sum of incident
    face coefficients + shift (SPD)."""
    n = k ** 3
    lo, hi, _ = _grid_edges(k, k, k, [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    rng = np.random.default_rng(seed)
    c = 10.0 ** rng.uniform(lo_exp, hi_exp, size=lo.size)
    diag = np.full(n, shift)
    np.add.at(diag, lo, c)
    np.add.at(diag, hi, c)
    rows = np.concatenate([np.arange(n), hi])
    cols = np.concatenate([np.arange(n), lo])
    vals = np.concatenate([diag, -c])
    return _lower_csc(n, rows, cols, vals)


_OFFS27 = [(dx, dy, dz) for dz in (0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
           if (dz, dy, dx) > (0, 0, 0)]


def stencil_3d_27pt(k=60, seed=0, lo_exp=-3.0, hi_exp=0.0, diag_factor=1.0 + 1e-6):
    """3D 27-point SPD stencil on a k^3 grid: for each of the 13 forward
    neighbour directions of every node, w = 10**U(lo_exp, hi_exp) (drawn in
    the edge order of the 13 directions), a_ij = a_ji = -w; the diagonal is
    the row absolute sum times diag_factor (strictly diagonally dominant)."""
    n = k ** 3
    lo, hi, _ = _grid_edges(k, k, k, _OFFS27)
    rng = np.random.default_rng(seed)
    w = 10.0 ** rng.uniform(lo_exp, hi_exp, size=lo.size)
    diag = np.zeros(n)
    np.add.at(diag, lo, w)
    np.add.at(diag, hi, w)
    diag *= diag_factor
    rows = np.concatenate([np.arange(n), hi])
    cols = np.concatenate([np.arange(n), lo])
    vals = np.concatenate([diag, -w])
    return _lower_csc(n, rows, cols, vals)


def shifted_laplacian_3d_7pt(k=80, shift=0.3):
    """A = L_h - shift*I with L_h the constant 7-point Laplacian (diag 6,
    off-diagonal -1, Dirichlet): symmetric indefinite for shift=0.3."""
    n = k ** 3
    lo, hi, _ = _grid_edges(k, k, k, [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    rows = np.concatenate([np.arange(n), hi])
    cols = np.concatenate([np.arange(n), lo])
    vals = np.concatenate([np.full(n, 6.0 - shift), np.full(lo.size, -1.0)])
    return _lower_csc(n, rows, cols, vals)


def random_laplacian(kind, k, seed):
    """Serving-mix request matrix: 2D 5-point ("2d", k x k) or 3D 7-point
    ("3d", k^3) Laplacian with face coefficients 10**U(-2,2)."""
    if kind == "3d":
        return diffusion_3d_7pt(k, seed=seed, shift=1e-3)
    n = k * k
    lo, hi, _ = _grid_edges(k, k, 1, [(1, 0, 0), (0, 1, 0)])
    rng = np.random.default_rng(seed)
    c = 10.0 ** rng.uniform(-2.0, 2.0, size=lo.size)
    diag = np.full(n, 1e-3)
    np.add.at(diag, lo, c)
    np.add.at(diag, hi, c)
    rows = np.concatenate([np.arange(n), hi])
    cols = np.concatenate([np.arange(n), lo])
    vals = np.concatenate([diag, -c])
    return _lower_csc(n, rows, cols, vals)


def random_spd(n, density, seed, diag_boost=1.0):
    """Random sparse SPD matrix (diagonally dominant) for structure tests."""
    rng = np.random.default_rng(seed)
    m = max(1, int(density * n * (n - 1) / 2))
    a = rng.integers(0, n, size=m)
    b = rng.integers(0, n, size=m)
    keep = a != b
    a, b = a[keep], b[keep]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = np.unique(lo * n + hi)
    lo, hi = key // n, key % n
    v = -rng.uniform(0.1, 1.0, size=lo.size)
    diag = np.full(n, diag_boost)
    np.add.at(diag, lo, -v)
    np.add.at(diag, hi, -v)
    rows = np.concatenate([np.arange(n), hi])
    cols = np.concatenate([np.arange(n), lo])
    vals = np.concatenate([diag, v])
    return _lower_csc(n, rows, cols, vals)


def full_from_lower(n, colptr, rowind, nzval):
    """Dense symmetric matrix from a lower CSC (small n only)."""
    A = np.zeros((n, n))
    for c in range(n):
        for p in range(colptr[c], colptr[c + 1]):
            r = rowind[p]
            A[r, c] = nzval[p]
            A[c, r] = nzval[p]
    return A
