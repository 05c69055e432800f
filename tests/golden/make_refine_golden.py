"""Writes tests/golden/refine.npz: orderings after the supernode reordering
(-refine SET10 .. SET31) as symPACK's own Fortran routine `ordsup` (with
`mcssup` and its heap routines) computes it, compiled from the reference
sources by oracle/Makefile into oracle/_ref/libref_orderings.so.

The routine's input is the unrefined supernodal structure (xsuper, xlindx,
lindx, supernode membership) of the library's symbolic factorization, whose
steps are themselves pinned bit-exact by the ordering and symbolic fixtures;
the expected ordering is the current one composed with the routine's output,
as refineSupernodes does (include/sympack/impl/symPACKMatrixBase_impl.hpp:
1006-1080, src/Ordering.cpp:1301-1314).

Run from the repository root (needs the reference tree and the compiled
library; the fixtures it writes are data only):
    make -C oracle ref && python3 tests/golden/make_refine_golden.py
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
CODES = ["SET10", "SET11", "SET20", "SET21", "SET30", "SET31"]


def graphs():
    """name -> (generator output, orderings); shared with tests/test_refine.py."""
    return {
        "lap2d_64": (M.laplacian_2d_5pt(64), ["MMD", "AMD"]),
        "diff3d_12": (M.diffusion_3d_7pt(12, seed=1), ["MMD", "AMD", "NDBOX"]),
        "st27_8": (M.stencil_3d_27pt(8, seed=2), ["MMD", "AMD", "NDBOX"]),
        "rand_1000": (M.random_spd(1000, 0.004, 2), ["MMD", "AMD"]),
        "diff3d_20": (M.diffusion_3d_7pt(20, seed=1), ["MMD"]),
    }


def symbolic(n, cp, ri, ordering, refine):
    A = sp.DistSparseMatrix(n, cp, ri, np.ones(len(ri)))
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering=ordering, order_refinement_str=refine))
    m.SymbolicFactorization(A)
    return m


def p32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def ref_refine(lib, m, ordflag, altflag):
    xs = m.GetSupernodalPartition().astype(np.int32)
    sn = m.GetSupMembership().astype(np.int32)
    xl, lx = m.GetStructure()
    xl = xl.astype(np.int32)
    lx = lx.astype(np.int32)
    n, ns = len(sn), len(xs) - 1
    perm = np.arange(1, n + 1, dtype=np.int32)
    invp = perm.copy()
    z = lambda k: np.zeros(k, np.int32)  # noqa: E731
    work = [z(n) for _ in range(10)]
    I = ctypes.c_int32
    lib.ordsup_(ctypes.byref(I(ordflag)), ctypes.byref(I(altflag)), ctypes.byref(I(n)),
                ctypes.byref(I(len(lx))), ctypes.byref(I(ns)), p32(xs), p32(xl), p32(lx), p32(sn),
                p32(perm), p32(invp), *[p32(w) for w in work[:9]], p32(work[9]), p32(z(ns)),
                p32(z(n + 1)), p32(z(n)), p32(z(n)), p32(z(n)), p32(z(2 * ns)))
    _, cur = m.GetOrdering()
    return invp[cur - 1].astype(np.int32)  # Ordering::Compose


def main():
    lib = ctypes.CDLL(LIB)
    out = {}
    for name, ((n, cp, ri, _), orderings) in graphs().items():
        for ordering in orderings:
            base = symbolic(n, cp, ri, ordering, "NO")
            for code in CODES:
                inv = ref_refine(lib, base, int(code[3]), int(code[4]))
                key = "%s_%s_%s" % (name, ordering, code)
                if n <= 1000:
                    out[key + "_invp"] = inv
                else:
                    out[key + "_sha256"] = np.frombuffer(hashlib.sha256(inv.tobytes()).digest(),
                                                         np.uint8)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "refine.npz"), **out)
    print("wrote %d arrays" % len(out))


if __name__ == "__main__":
    main()
