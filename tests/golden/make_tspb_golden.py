"""Writes tests/golden/tspb.npz: orderings after the traveling-salesman
supernode reordering (-refine TSPB, and TSPB1 for the other TSPB* names) as
symPACK's own Fortran routines compute them (src/TSPB/
ordsup_ind_tsp_paths2.f with skeleton compression, ordsup_ind_tsp_paths.f
without, and the routines they call), compiled from the reference sources by
oracle/Makefile into oracle/_ref/libref_orderings.so.

The routines' inputs are what symPACKMatrix2D passes them
(include/sympack/symPACKMatrix2D.hpp:3113-3243): the expanded pattern of A
without its diagonal in the original numbering (1-based), the postordered
elimination-tree parents, the current ordering, and the unrefined
supernodal structure (xsuper, xlindx, lindx, supernode membership) of the
library's symbolic factorization, whose steps are pinned bit-exact by the
ordering and symbolic fixtures.  The routines compose the new ordering into
the current one themselves (invinv.f); the fixture is the resulting invp.

Run from the repository root (needs the reference tree and the compiled
library; the fixtures it writes are data only):
    make -C oracle ref && python3 tests/golden/make_tspb_golden.py
"""
import ctypes
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import sympack_amd as sp  # noqa: E402
from sympack_amd import matrices as M  # noqa: E402

LIB = os.path.join(ROOT, "oracle", "_ref", "libref_orderings.so")
CODES = ["TSPB", "TSPB1"]


def graphs():
    """name -> (generator output, [(ordering, maximum supernode width)]);
    shared with tests/test_refine.py."""
    return {
        "lap2d_64": (M.laplacian_2d_5pt(64), [("MMD", 0), ("AMD", 0), ("NDGRID", 64)]),
        "diff3d_12": (M.diffusion_3d_7pt(12, seed=1), [("MMD", 0), ("AMD", 0), ("NDBOX", 40)]),
        "st27_8": (M.stencil_3d_27pt(8, seed=2), [("MMD", 0), ("NDBOX", 60)]),
        "rand_1000": (M.random_spd(1000, 0.004, 2), [("MMD", 0), ("AMD", 0)]),
        "diff3d_20": (M.diffusion_3d_7pt(20, seed=1), [("MMD", 0), ("NDBOX", 150)]),
    }


def symbolic(n, cp, ri, ordering, maxsize, refine):
    A = sp.DistSparseMatrix(n, cp, ri, np.ones(len(ri)))
    rp = sp.RelaxationParameters()
    if maxsize:
        rp.SetMaxSize(maxsize)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering=ordering, relax=rp, order_refinement_str=refine))
    m.SymbolicFactorization(A)
    return m


def expanded_graph(n, cp, ri):
    """Both triangles of the lower CSC pattern (0-based cp/ri), diagonal
    dropped, columns sorted; 1-based xadj / adjncy."""
    rows, cols = [], []
    for c in range(n):
        for q in range(cp[c], cp[c + 1]):
            r = int(ri[q])
            if r != c:
                rows += [r, c]
                cols += [c, r]
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    order = np.lexsort((rows, cols))
    rows, cols = rows[order], cols[order]
    xadj = np.zeros(n + 1, np.int32)
    np.add.at(xadj, cols + 1, 1)
    xadj = (np.cumsum(xadj) + 1).astype(np.int32)
    return xadj, (rows + 1).astype(np.int32)


def p32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def ref_tspb(lib, m, n, cp, ri, compress):
    xadj, adj = expanded_graph(n, cp, ri)
    xs = m.GetSupernodalPartition().astype(np.int32)
    sn = m.GetSupMembership().astype(np.int32)
    xl, lx = m.GetStructure()
    xl = xl.astype(np.int32)
    lx = lx.astype(np.int32)
    perm, invp = m.GetOrdering()
    perm = perm.astype(np.int32).copy()
    invp = invp.astype(np.int32).copy()
    etpar = m.GetETree().astype(np.int32)
    ns = len(xs) - 1
    supsiz = int(np.max(np.diff(xs)))
    nadj = len(adj)
    iwsiz = 8 * n + 3
    z = lambda k: np.zeros(k, np.int32)  # noqa: E731
    I = ctypes.c_int32
    iflag = I(0)
    args = [ctypes.byref(I(nadj)), ctypes.byref(I(n)), ctypes.byref(I(len(lx))), ctypes.byref(I(ns)),
            ctypes.byref(I(supsiz)), p32(xs), p32(xl), p32(lx), p32(sn), p32(xadj), p32(adj), p32(etpar),
            p32(perm), p32(invp), ctypes.byref(iflag),
            p32(z(n + 1)), p32(z(n)), p32(z(n)), p32(z(nadj)), p32(z(n)), p32(z(n)), p32(z(ns)), p32(z(ns)),
            p32(z(n)), p32(z((supsiz + 1) * (supsiz + 1))), p32(z(ns)), ctypes.byref(I(iwsiz)), p32(z(iwsiz))]
    if compress:
        lib.ordsup_ind_tsp_paths2_(*args, p32(z(n)))
    else:
        lib.ordsup_ind_tsp_paths_(*args)
    assert iflag.value == 0
    return invp


def key(name, ordering, maxsize, code):
    return "%s_%s%s_%s" % (name, ordering, maxsize or "", code)


def main():
    lib = ctypes.CDLL(LIB)
    out = {}
    for name, ((n, cp, ri, _), cases) in graphs().items():
        for ordering, maxsize in cases:
            base = symbolic(n, cp, ri, ordering, maxsize, "NO")
            for code in CODES:
                inv = ref_tspb(lib, base, n, cp, ri, code == "TSPB")
                k = key(name, ordering, maxsize, code)
                if n <= 1000:
                    out[k + "_invp"] = inv
                else:
                    out[k + "_sha256"] = np.frombuffer(hashlib.sha256(inv.tobytes()).digest(), np.uint8)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "tspb.npz"), **out)
    print("wrote %d arrays" % len(out))


if __name__ == "__main__":
    main()
