"""Host checks of the Schur-update merging the plan chooses: a supernode I
deferred into P splits its update into the columns of P (applied at I's
level) and the rest, which P's own update computes as extra k columns.  That
is only valid when P is I's etree parent and every row of struct(I) beyond
P's columns lies in struct(P); the plan also keeps each merge to one child,
never defers a supernode that absorbs one, and requires the deferred rows to
cover most of P's off-diagonal rows.  The decisive check replays the merged
updates on the dense pattern: every (source supernode, target entry) pair of
the unmerged right-looking algorithm is computed exactly once."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M


def _plan(gen, maxsize, merge=2):
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(maxsize)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, merge_updates=merge))
    m.SymbolicFactorization(A)
    return m


CASES = [
    ("lap3d_16_w16", lambda: M.diffusion_3d_7pt(16, seed=1), 16),
    ("lap3d_20_w24", lambda: M.diffusion_3d_7pt(20, seed=2), 24),
    ("st27_12_w20", lambda: M.stencil_3d_27pt(12, seed=3), 20),
    ("lap2d_64_w8", lambda: M.laplacian_2d_5pt(64), 8),
]


@pytest.mark.parametrize("name,gen,maxsize", CASES, ids=[c[0] for c in CASES])
def test_merge_rules(name, gen, maxsize):
    m = _plan(gen, maxsize)
    d = m.GetUpdateMerges()
    xs = np.asarray(m.GetSupernodalPartition()) - 1
    xl, lx = m.GetStructure()
    xl = xl - 1
    lx = lx - 1
    ns = len(xs) - 1
    sup = np.zeros(xs[-1], np.int64)
    for I in range(ns):
        sup[xs[I]:xs[I + 1]] = I
    merged = d[d >= 0]
    assert len(merged) > 0, "expected chains to merge in %s" % name
    assert len(set(merged.tolist())) == len(merged), "a parent absorbs at most one child"
    for I in np.nonzero(d >= 0)[0]:
        P = int(d[I])
        assert d[P] == -1, "a supernode that absorbs a child is not deferred"
        w = xs[I + 1] - xs[I]
        rows = lx[xl[I] + w:xl[I + 1]]
        assert sup[rows[0]] == P, "deferred into the etree parent"
        n1 = int(np.sum(rows < xs[P + 1]))
        r2 = rows[n1:]
        wp = xs[P + 1] - xs[P]
        prow = lx[xl[P] + wp:xl[P + 1]]
        assert np.all(np.isin(r2, prow)), "deferred rows lie in struct(P)"
        assert len(r2) >= 32 and len(r2) >= 0.85 * len(prow)
    # replay: supernode X's tasks compute, for the pairs (r1 >= r2) of X's
    # off-diagonal rows with r2 below the cut (a deferred X: its parent's
    # columns), the contributions of X itself and of the child X absorbed
    # (rows the child lacks contribute nothing)
    W = np.diff(xs)
    rows_of = [lx[xl[I] + W[I]:xl[I + 1]] for I in range(ns)]
    child_of = {int(P): int(I) for I, P in enumerate(d) if P >= 0}
    count = {}
    for X in range(ns):
        rows = rows_of[X]
        if len(rows) == 0:
            continue
        cut_end = xs[d[X] + 1] if d[X] >= 0 else np.iinfo(np.int64).max
        for D in [X] + ([child_of[X]] if X in child_of else []):
            have = set(rows_of[D].tolist())
            rr = [r for r in rows.tolist() if r in have]
            for i, r2 in enumerate(rr):
                if r2 >= cut_end:
                    break
                for r1 in rr[i:]:
                    count[(D, r1, r2)] = count.get((D, r1, r2), 0) + 1
    expected = {(D, r1, r2) for D in range(ns) for i, r2 in enumerate(rows_of[D].tolist())
                for r1 in rows_of[D][i:].tolist()}
    assert set(count) == expected, "every unmerged update entry is computed"
    assert all(v == 1 for v in count.values()), "and only once"
    # off: nothing deferred
    assert np.all(_plan(gen, maxsize, merge=0).GetUpdateMerges() == -1)


def test_auto_merging_only_on_wide_levels():
    """Default mode: merges only where both levels hold >= 768 update tiles
    (3 per CU), so none on a small matrix and none near the root."""
    assert np.all(_plan(lambda: M.diffusion_3d_7pt(16, seed=1), 16, merge=1).GetUpdateMerges() == -1)
    m = _plan(lambda: M.diffusion_3d_7pt(48, seed=0), 150, merge=1)
    d = m.GetUpdateMerges()
    xs = np.asarray(m.GetSupernodalPartition()) - 1
    xl, lx = m.GetStructure()
    xl, lx = xl - 1, lx - 1
    ns = len(xs) - 1
    sup = np.zeros(xs[-1], np.int64)
    for I in range(ns):
        sup[xs[I]:xs[I + 1]] = I
    W = np.diff(xs)
    Mr = np.diff(xl) - W
    lev = np.zeros(ns, np.int64)
    for I in range(ns):
        if Mr[I] > 0:
            P = sup[lx[xl[I] + W[I]]]
            lev[P] = max(lev[P], lev[I] + 1)
    T = -(-Mr // 64)
    tiles = np.bincount(lev, weights=T * (T + 1) // 2)
    merged = np.nonzero(d >= 0)[0]
    assert len(merged) > 0
    for I in merged:
        assert tiles[lev[I]] >= 768 and tiles[lev[d[I]]] >= 768
    # the top of the tree (levels with a single supernode) merges nothing
    top = lev.max()
    assert np.all(d[lev >= top - 5] == -1)


@pytest.mark.parametrize("name,gen,maxsize", CASES + [("diff3d_48_w150", lambda: M.diffusion_3d_7pt(48, seed=0), 150)],
                         ids=[c[0] for c in CASES] + ["diff3d_48_w150"])
def test_dependency_driven_solve_plan(name, gen, maxsize):
    """The solve's task lists: in ticket order no wait blocks, every counter
    drains and every tile's targets are the owners of its rows."""
    _plan(gen, maxsize).ValidateSolvePlan()
