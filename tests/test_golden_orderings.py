"""Orderings against golden vectors written by symPACK's own Fortran routines
(ordmmd/genmmd, genrcm, boxnd, gridnd) compiled from the reference sources
(tests/golden/make_golden.py).  Bit-exact."""
import hashlib
import os

import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import symbolic as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "orderings.npz"))
GRAPHS = sorted({k[:-7] for k in G.files if k.endswith("_colptr")})


def _A(name):
    cp = G[name + "_colptr"]
    ri = G[name + "_rowind"]
    n = len(cp) - 1
    return sp.DistSparseMatrix(n, cp, ri, np.ones(len(ri)), baseval=0)


@pytest.mark.parametrize("name", GRAPHS)
def test_mmd_matches_reference_fortran(name):
    perm, invp = sp.compute_ordering("MMD", _A(name))
    assert np.array_equal(invp, G[name + "_mmd_invp"])
    assert np.array_equal(np.argsort(invp) + 1, perm)


@pytest.mark.parametrize("name", GRAPHS)
def test_amd_matches_reference_fortran(name):
    perm, invp = sp.compute_ordering("AMD", _A(name))
    assert np.array_equal(invp, G[name + "_amd_invp"])
    assert np.array_equal(np.argsort(invp) + 1, perm)


@pytest.mark.parametrize("name,gen", [("st27_20", lambda: M.stencil_3d_27pt(20, seed=4)),
                                      ("diff3d_30", lambda: M.diffusion_3d_7pt(30, seed=5))])
def test_amd_large_checksum(name, gen):
    n, cp, ri, v = gen()
    perm, invp = sp.compute_ordering("AMD", sp.DistSparseMatrix(n, cp, ri, v))
    digest = hashlib.sha256(invp.astype(np.int32).tobytes()).digest()
    assert digest == G["amd_%s_sha256" % name].tobytes()


@pytest.mark.parametrize("name", GRAPHS)
def test_rcm_matches_reference_fortran(name):
    perm, invp = sp.compute_ordering("RCM", _A(name))
    assert np.array_equal(perm, G[name + "_rcm_perm"])


@pytest.mark.parametrize("k", [4, 6, 9, 12, 20])
def test_ndbox_matches_reference_fortran(k):
    n = k ** 3
    A = sp.DistSparseMatrix(n, np.arange(n + 1), np.arange(n), np.ones(n), baseval=0)
    perm, invp = sp.compute_ordering("NDBOX", A)
    assert np.array_equal(invp, G["boxnd_%d" % k])
    assert np.array_equal(O.boxnd(k, k, k), G["boxnd_%d" % k])  # the Python restatement too


@pytest.mark.parametrize("k", [48, 60])
def test_ndbox_large_checksum(k):
    n = k ** 3
    A = sp.DistSparseMatrix(n, np.arange(n + 1), np.arange(n), np.ones(n), baseval=0)
    perm, invp = sp.compute_ordering("NDBOX", A)
    assert hashlib.sha256(invp.astype(np.int32).tobytes()).digest() == G["boxnd_%d_sha256" % k].tobytes()


@pytest.mark.parametrize("k", [5, 8, 40, 64])
def test_ndgrid_matches_reference_fortran(k):
    n = k * k
    A = sp.DistSparseMatrix(n, np.arange(n + 1), np.arange(n), np.ones(n), baseval=0)
    perm, invp = sp.compute_ordering("NDGRID", A)
    assert np.array_equal(invp, G["gridnd_%d" % k])
    assert np.array_equal(O.gridnd(k, k), G["gridnd_%d" % k])


def test_ndbox_rejects_non_cube():
    n = 1000 + 1
    A = sp.DistSparseMatrix(n, np.arange(n + 1), np.arange(n), np.ones(n), baseval=0)
    with pytest.raises(ValueError):
        sp.compute_ordering("NDBOX", A)
