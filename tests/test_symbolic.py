"""Symbolic analysis parity: the C++ chain (etree, postorder, SortChildren,
column counts, supernodes, relaxation, structure, cell map) against the
Python restatement (oracle/symbolic.py) and brute-force symbolic
elimination, starting from orderings pinned by the reference Fortran."""
import os

import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import symbolic as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "orderings.npz"))


def _analyze(A, ordering, relax=None, perm=None):
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering=ordering, relax=relax or sp.RelaxationParameters(), perm=perm))
    m.SymbolicFactorization(A)
    return m


def _check_chain(A, m, perm0, relax=None):
    r = O.symbolic_chain(A.size, A.colptr - A.baseval, A.rowind - A.baseval, list(perm0), relax=relax)
    perm, invp = m.GetOrdering()
    assert list(perm) == r["perm"]
    assert list(m.GetETree()) == r["parent"]
    ccu, ccr = m.GetColumnCounts()
    assert list(ccu) == r["cc_unrelaxed"]
    assert list(ccr) == r["cc"]
    assert list(m.GetSupernodalPartition()) == r["xsuper"]
    assert list(m.GetSupMembership()) == r["sup"]
    xl, lx = m.GetStructure()
    for I in range(len(r["xsuper"]) - 1):
        assert list(lx[xl[I] - 1:xl[I + 1] - 1]) == r["structs"][I]
    st = m.stats()
    assert st["flops"] == r["flops"]
    assert st["nnz_l"] == r["nnzL"]
    return r


CASES = [
    ("lap2d_64", "MMD"), ("lap2d_17", "MMD"), ("diff3d_10", "MMD"), ("st27_8", "MMD"),
    ("rand_300", "MMD"), ("rand_1000", "MMD"), ("rand_2000_sparse", "MMD"),
    ("lap2d_64", "RCM"), ("diff3d_10", "RCM"), ("rand_1000", "RCM"),
    ("lap2d_64", "AMD"), ("st27_8", "AMD"), ("rand_2000_sparse", "AMD"),
]


@pytest.mark.parametrize("name,ordering", CASES)
def test_chain_from_reference_ordering(name, ordering):
    cp = G[name + "_colptr"]
    ri = G[name + "_rowind"]
    n = len(cp) - 1
    A = sp.DistSparseMatrix(n, cp, ri, np.ones(len(ri)), baseval=0)
    if ordering in ("MMD", "AMD"):
        invp = G[name + "_%s_invp" % ordering.lower()]
        perm0 = np.argsort(invp) + 1
    else:
        perm0 = G[name + "_rcm_perm"]
    m = _analyze(A, ordering)
    _check_chain(A, m, perm0)


@pytest.mark.parametrize("k,maxsize", [(6, 0), (8, 0), (12, 0), (12, 150), (12, 20), (16, 150)])
def test_chain_ndbox(k, maxsize):
    n, cp, ri, v = M.diffusion_3d_7pt(k, seed=k)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(maxsize)
    m = _analyze(A, "NDBOX", rp)
    _check_chain(A, m, O.invp_to_perm(O.boxnd(k, k, k)),
                 relax=dict(max_size=maxsize, nrelax0=rp.nrelax0, nrelax1=rp.nrelax1, nrelax2=rp.nrelax2))


@pytest.mark.parametrize("relax", [dict(nrelax0=0), dict(nrelax0=4, nrelax1=16, nrelax2=32),
                                   dict(zrelax0=0.5, zrelax1=0.2, zrelax2=0.1)])
def test_chain_relaxation_parameters(relax):
    n, cp, ri, v = M.laplacian_2d_5pt(30)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters(**{k: v for k, v in relax.items()})
    m = _analyze(A, "NDGRID", rp)
    r = dict(nrelax0=rp.nrelax0, nrelax1=rp.nrelax1, nrelax2=rp.nrelax2,
             zrelax0=rp.zrelax0, zrelax1=rp.zrelax1, zrelax2=rp.zrelax2)
    _check_chain(A, m, O.invp_to_perm(O.gridnd(30, 30)), relax=r)


def test_etree_and_counts_brute_force():
    n, cp, ri, v = M.random_spd(500, 0.01, 11)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    m = _analyze(A, "MMD")
    perm, invp = m.GetOrdering()
    adj = O.expand_symmetric(n, cp, ri)
    cc, parent = O.column_counts(n, adj, list(perm), list(invp))
    assert list(m.GetETree()) == parent
    assert list(m.GetColumnCounts()[0]) == cc


def test_natural_and_user_orderings():
    n, cp, ri, v = M.laplacian_2d_5pt(20)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    m = _analyze(A, "NATURAL")
    # NATURAL skips SortChildren (symPACKMatrix2D.hpp:2929)
    r = O.symbolic_chain(n, cp, ri, list(range(1, n + 1)), ordering_kind="natural")
    assert list(m.GetOrdering()[0]) == r["perm"]
    assert list(m.GetSupernodalPartition()) == r["xsuper"]
    rng = np.random.default_rng(3)
    up = rng.permutation(n) + 1
    m2 = _analyze(A, "USER", perm=up)
    r2 = O.symbolic_chain(n, cp, ri, list(up))
    assert list(m2.GetOrdering()[0]) == r2["perm"]
    assert list(m2.GetSupernodalPartition()) == r2["xsuper"]


def test_user_ordering_requires_perm():
    n, cp, ri, v = M.laplacian_2d_5pt(8)
    m = sp.symPACKMatrix2D()
    with pytest.raises(ValueError):
        m.Init(sp.symPACKOptions(ordering="USER"))


def test_bad_ordering_string():
    with pytest.raises(ValueError):
        sp.symPACKOptions.from_args(["-ordering", "METIS"])


def test_cell_map_invariants():
    n, cp, ri, v = M.laplacian_2d_5pt(32)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    m = _analyze(A, "MMD")
    cells = m.GetCells()
    xs = m.GetSupernodalPartition()
    sup = m.GetSupMembership()
    xl, lx = m.GetStructure()
    nsuper = len(xs) - 1
    # cells appear for I descending; per I, K ascending, K >= I, rows partition struct(I)
    b = 0
    seen = {}
    for c in range(len(cells["K"])):
        K, I = cells["K"][c], cells["I"][c]
        assert K >= I
        rows = []
        for q in range(cells["nblocks"][c]):
            fr, nr = cells["block_first_row"][b], cells["block_nrows"][b]
            b += 1
            rows += list(range(fr, fr + nr))
        assert len(rows) == cells["nrows"][c]
        assert all(sup[r - 1] == K for r in rows)
        seen.setdefault(I, []).extend(rows)
    assert b == len(cells["block_first_row"])
    for I in range(1, nsuper + 1):
        assert seen[I] == list(lx[xl[I - 1] - 1:xl[I] - 1])
    Is = list(cells["I"])
    assert Is == sorted(Is, reverse=True)
