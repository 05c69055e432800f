"""Supernode reordering (-refine SET10 .. SET31; symPACKMatrix2D.hpp:3090-3111,
refineSupernodes in include/sympack/impl/symPACKMatrixBase_impl.hpp:
1006-1080, src/ordsup.f, src/mcssup.f) against golden orderings written by
the reference's own Fortran routines (tests/golden/make_refine_golden.py).
Bit-exact."""
import hashlib
import os
import sys

import numpy as np
import pytest

import sympack_amd as sp

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_refine_golden import CODES, graphs, symbolic  # noqa: E402

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "refine.npz"))
GRAPHS = graphs()
CASES = [(name, o, c) for name, (_, ords) in GRAPHS.items() for o in ords for c in CODES]


@pytest.mark.parametrize("name,ordering,code", CASES,
                         ids=["%s-%s-%s" % c for c in CASES])
def test_refine_matches_reference_fortran(name, ordering, code):
    n, cp, ri, _ = GRAPHS[name][0]
    m = symbolic(n, cp, ri, ordering, code)
    perm, invp = m.GetOrdering()
    key = "%s_%s_%s" % (name, ordering, code)
    if key + "_invp" in G.files:
        assert np.array_equal(invp, G[key + "_invp"])
    else:
        digest = hashlib.sha256(invp.astype(np.int32).tobytes()).digest()
        assert digest == G[key + "_sha256"].tobytes()
    assert np.array_equal(np.argsort(invp) + 1, perm)


@pytest.mark.parametrize("name,ordering", [("diff3d_12", "MMD"), ("st27_8", "AMD"),
                                           ("rand_1000", "MMD")])
def test_refine_keeps_partition_and_cuts_row_blocks(name, ordering):
    n, cp, ri, _ = GRAPHS[name][0]
    base = symbolic(n, cp, ri, ordering, "NO")
    ref = symbolic(n, cp, ri, ordering, "SET")
    xs = base.GetSupernodalPartition()
    assert np.array_equal(ref.GetSupernodalPartition(), xs)
    p0, _ = base.GetOrdering()
    p1, _ = ref.GetOrdering()
    assert not np.array_equal(p0, p1)
    # only the numbering inside each supernode changes
    for s in range(len(xs) - 1):
        a, b = xs[s] - 1, xs[s + 1] - 1
        assert np.array_equal(np.sort(p0[a:b]), np.sort(p1[a:b]))
    assert len(ref.GetStructure()[1]) == len(base.GetStructure()[1])
    assert ref.stats()["flops"] == base.stats()["flops"]
    # what the reordering is for: fewer runs of consecutive rows (blocks)
    assert len(ref.GetCells()["block_nrows"]) < len(base.GetCells()["block_nrows"])


def test_refine_option_parsing():
    o = sp.symPACKOptions.from_args(["-ordering", "AMD", "-refine", "SET"])
    assert o.order_refinement_str == "SET31"
    assert sp.symPACKOptions.from_args(["-refine", "SET20"]).order_refinement_str == "SET20"
    assert sp.symPACKOptions.from_args(["-refine", "NO"]).order_refinement_str == "NO"
    assert sp.symPACKOptions.from_args([]).order_refinement_str == "NO"
    assert sp.symPACKOptions.from_args(["-refine", "TSPB"]).order_refinement_str == "TSPB"
    assert sp.symPACKOptions.from_args(["-refine", "TSPB2"]).to_c().refine == 41
    assert sp.symPACKOptions(order_refinement_str="TSPB").to_c().refine == 40
    assert sp.symPACKOptions(order_refinement_str="TSPBX").to_c().refine == 41


def test_unknown_refinement_name_leaves_ordering():
    # names other than SET* and TSPB* only rerun the structure (the reference
    # compares the string against each name in turn)
    n, cp, ri, _ = GRAPHS["st27_8"][0]
    a = symbolic(n, cp, ri, "MMD", "NO").GetOrdering()[1]
    b = symbolic(n, cp, ri, "MMD", "SET99").GetOrdering()[1]
    assert np.array_equal(a, b)


@pytest.mark.parametrize("code", ["SET", "SET10", "SET20"])
def test_refined_structure_factors_on_cpu(code):
    # the refined numbering still has all fill inside the supernodal
    # structure: the oracle's factor on it solves the system
    from oracle import numeric as onum
    n, cp, ri, v = GRAPHS["st27_8"][0]
    A = sp.DistSparseMatrix(n, cp, ri, v)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="MMD", order_refinement_str=code))
    m.SymbolicFactorization(A)
    xt = np.random.default_rng(0).uniform(-1, 1, n)
    b = A.matvec(xt)
    x = onum.factor_and_solve(m, A, b)
    assert onum.backward_error(A, x, b) < 1e-14
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-10


# -refine TSPB* (symPACKMatrix2D.hpp:3113-3263, src/TSPB/*.f) against the
# reference's own Fortran (tests/golden/make_tspb_golden.py): TSPB runs the
# tour with skeleton compression (ordsup_ind_tsp_paths2), any other TSPB*
# name the tour without it (ordsup_ind_tsp_paths).  Bit-exact.
sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_tspb_golden as TG  # noqa: E402

GT = np.load(os.path.join(os.path.dirname(__file__), "golden", "tspb.npz"))
TGRAPHS = TG.graphs()
TCASES = [(name, o, ms, c) for name, (_, cases) in TGRAPHS.items() for o, ms in cases for c in TG.CODES]


@pytest.mark.parametrize("name,ordering,maxsize,code", TCASES, ids=["%s-%s%s-%s" % c for c in TCASES])
def test_tspb_matches_reference_fortran(name, ordering, maxsize, code):
    n, cp, ri, _ = TGRAPHS[name][0]
    m = TG.symbolic(n, cp, ri, ordering, maxsize, code)
    perm, invp = m.GetOrdering()
    k = TG.key(name, ordering, maxsize, code)
    if k + "_invp" in GT.files:
        assert np.array_equal(invp, GT[k + "_invp"])
    else:
        digest = hashlib.sha256(invp.astype(np.int32).tobytes()).digest()
        assert digest == GT[k + "_sha256"].tobytes()
    assert np.array_equal(np.argsort(invp) + 1, perm)


@pytest.mark.parametrize("code", ["TSPB", "TSPB1"])
def test_tspb_keeps_partition_and_counts(code):
    n, cp, ri, _ = TGRAPHS["diff3d_12"][0]
    base = TG.symbolic(n, cp, ri, "MMD", 0, "NO")
    ref = TG.symbolic(n, cp, ri, "MMD", 0, code)
    xs = base.GetSupernodalPartition()
    assert np.array_equal(ref.GetSupernodalPartition(), xs)
    p0, _ = base.GetOrdering()
    p1, _ = ref.GetOrdering()
    assert not np.array_equal(p0, p1)
    for s in range(len(xs) - 1):  # only the numbering inside each supernode changes
        a, b = xs[s] - 1, xs[s + 1] - 1
        assert np.array_equal(np.sort(p0[a:b]), np.sort(p1[a:b]))
    assert len(ref.GetStructure()[1]) == len(base.GetStructure()[1])
    assert ref.stats()["flops"] == base.stats()["flops"]


def test_tspb_structure_factors_on_cpu():
    from oracle import numeric as onum
    n, cp, ri, v = TGRAPHS["st27_8"][0]
    A = sp.DistSparseMatrix(n, cp, ri, v)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="MMD", order_refinement_str="TSPB"))
    m.SymbolicFactorization(A)
    xt = np.random.default_rng(0).uniform(-1, 1, n)
    b = A.matvec(xt)
    x = onum.factor_and_solve(m, A, b)
    assert onum.backward_error(A, x, b) < 1e-14
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-10
