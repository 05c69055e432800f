"""Writes tests/golden/orderings.npz: ordering vectors produced by symPACK's
own Fortran routines (ordmmd/genmmd, amdbar, genrcm, boxnd, gridnd), compiled from the
reference sources by oracle/Makefile into oracle/_ref/libref_orderings.so.

Run from the repository root (needs the reference tree and the compiled
library; the fixtures it writes are data only):
    make -C oracle ref && python3 tests/golden/make_golden.py
Inputs are the repository's own generators (sympack_amd/matrices.py) with the
seeds recorded below, plus the CSC of the random graphs themselves.
"""
import ctypes
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sympack_amd import matrices as M  # noqa: E402

LIB = os.path.join(ROOT, "oracle", "_ref", "libref_orderings.so")


def p32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def graph_nodiag(n, cp, ri):
    """Expanded adjacency without diagonal, 1-based (xadj, adjncy), columns
    sorted: the graph Ordering::MMD/RCM hand to Fortran."""
    rows, cols = [], []
    for c in range(n):
        for p in range(cp[c], cp[c + 1]):
            r = int(ri[p])
            if r != c:
                rows += [r, c]
                cols += [c, r]
    rows = np.array(rows, np.int64)
    cols = np.array(cols, np.int64)
    order = np.lexsort((rows, cols))
    rows, cols = rows[order], cols[order]
    xadj = np.zeros(n + 1, np.int32)
    np.add.at(xadj, cols + 1, 1)
    xadj = (np.cumsum(xadj) + 1).astype(np.int32)
    return xadj, (rows + 1).astype(np.int32)


def ref_mmd(lib, n, cp, ri):
    xadj, adjncy = graph_nodiag(n, cp, ri)
    invp = np.zeros(n, np.int32)
    perm = np.zeros(n, np.int32)
    iw = np.zeros(4 * n, np.int32)
    I = ctypes.c_int32
    lib.ordmmd_(ctypes.byref(I(n)), ctypes.byref(I(len(adjncy))), p32(xadj), p32(adjncy.copy()),
                p32(invp), p32(perm), ctypes.byref(I(4 * n)), p32(iw), ctypes.byref(I(0)),
                ctypes.byref(I(0)))
    return invp  # Ordering::MMD keeps invp and rebuilds perm from it


def ref_amd(lib, n, cp, ri):
    """amdbar_ called with the arrays Ordering::AMD builds (Ordering.cpp:
    456-486): IWLEN = nnz + n, PFREE = nnz + 1; ELEN returns invp."""
    xadj, adjncy = graph_nodiag(n, cp, ri)
    nnz = len(adjncy)
    pe = xadj.copy()
    iw = np.zeros(nnz + n, np.int32)
    iw[:nnz] = adjncy
    ln = np.diff(xadj).astype(np.int32)
    nv, nxt, last, head, elen, deg, w = (np.zeros(n, np.int32) for _ in range(7))
    I = ctypes.c_int32
    ncmpa = I(0)
    lib.amdbar_(ctypes.byref(I(n)), p32(pe), p32(iw), p32(ln), ctypes.byref(I(nnz + n)),
                ctypes.byref(I(nnz + 1)), p32(nv), p32(nxt), p32(last), p32(head), p32(elen), p32(deg),
                ctypes.byref(ncmpa), p32(w), ctypes.byref(I(2 ** 31 - 1)))
    return elen


def ref_rcm(lib, n, cp, ri):
    xadj, adjncy = graph_nodiag(n, cp, ri)
    perm = np.zeros(2 * n, np.int32)
    mask = np.zeros(2 * n, np.int32)
    xls = np.zeros(2 * n, np.int32)
    lib.genrcm_(ctypes.byref(ctypes.c_int32(n)), p32(xadj), p32(adjncy.copy()), p32(perm), p32(mask),
                p32(xls))
    return perm[:n].copy()


def ref_boxnd(lib, k):
    n = k ** 3
    box = np.zeros(n, np.int32)
    stack = np.zeros(n if k * k > 25 else 2 * n, np.int32)
    I = ctypes.c_int32
    lib.boxnd_(ctypes.byref(I(k)), ctypes.byref(I(k)), ctypes.byref(I(k)), p32(box), p32(stack),
               ctypes.byref(I(len(stack))), ctypes.byref(I(1)))
    return box


def ref_gridnd(lib, k):
    n = k * k
    grid = np.zeros(n, np.int32)
    stack = np.zeros(n if k * k > 25 else 2 * n, np.int32)
    I = ctypes.c_int32
    lib.gridnd_(ctypes.byref(I(k)), ctypes.byref(I(k)), p32(grid), p32(stack),
                ctypes.byref(I(len(stack))), ctypes.byref(I(1)))
    return grid


def main():
    lib = ctypes.CDLL(LIB)
    out = {}
    graphs = {
        "lap2d_64": M.laplacian_2d_5pt(64),
        "lap2d_17": M.laplacian_2d_5pt(17),
        "diff3d_10": M.diffusion_3d_7pt(10, seed=1),
        "st27_8": M.stencil_3d_27pt(8, seed=2),
        "rand_300": M.random_spd(300, 0.02, 1),
        "rand_1000": M.random_spd(1000, 0.004, 2),
        "rand_2000_sparse": M.random_spd(2000, 0.0008, 3),
    }
    for name, (n, cp, ri, v) in graphs.items():
        out[name + "_colptr"] = cp.astype(np.int64)
        out[name + "_rowind"] = ri.astype(np.int32)
        out[name + "_mmd_invp"] = ref_mmd(lib, n, cp, ri)
        out[name + "_amd_invp"] = ref_amd(lib, n, cp, ri)
        out[name + "_rcm_perm"] = ref_rcm(lib, n, cp, ri)
    for k in (4, 6, 9, 12, 20):
        out["boxnd_%d" % k] = ref_boxnd(lib, k)
    for k in (5, 8, 40, 64):
        out["gridnd_%d" % k] = ref_gridnd(lib, k)
    # large boxes: checksum only
    for k in (48, 60):
        box = ref_boxnd(lib, k)
        out["boxnd_%d_sha256" % k] = np.frombuffer(hashlib.sha256(box.tobytes()).digest(), np.uint8)
    # larger AMD cases (workspace compaction): checksum of invp only
    for name, (n, cp, ri, v) in {"st27_20": M.stencil_3d_27pt(20, seed=4),
                                 "diff3d_30": M.diffusion_3d_7pt(30, seed=5)}.items():
        invp = ref_amd(lib, n, cp, ri)
        out["amd_%s_sha256" % name] = np.frombuffer(hashlib.sha256(invp.tobytes()).digest(), np.uint8)
    path = os.path.join(ROOT, "tests", "golden", "orderings.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
