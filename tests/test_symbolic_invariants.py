"""Independent pins of the symbolic chain: postorder, SortChildren,
fundamental supernodes, relaxation and supernodal structure are checked
against properties fixed by their definitions and against a dense boolean
symbolic elimination written here (not the oracle's set-based one).  Every
check runs on both the C++ product chain and the Python oracle
(oracle/symbolic.py), so a misreading shared by the two restatements of
ETree.cpp:59-225 and symPACKMatrixBase_impl.hpp:42-609 fails here.

The relaxation's merge decisions (symPACKMatrixBase_impl.hpp:473-609) are
replayed on the dense pattern: the zeros each merge introduces are counted
entry by entry (and must equal the reference's closed form), the column
counts come from the pattern, and the resulting partition must be the
product's.  The comparison rule itself (thresholds by merged width) is the
reference's, restated."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import symbolic as O


# ----------------------------------------------------------------- helpers
def dense_pattern(n, colptr, rowind, perm):
    """Boolean pattern of P A P^T (full symmetric), perm[new-1] = old."""
    invp = np.empty(n, np.int64)
    invp[np.asarray(perm) - 1] = np.arange(n)
    P = np.zeros((n, n), bool)
    for c in range(n):
        for q in range(colptr[c], colptr[c + 1]):
            r = int(rowind[q])
            a, b = invp[r], invp[c]
            P[a, b] = P[b, a] = True
    np.fill_diagonal(P, True)
    return P


def eliminate(P):
    """Lower pattern of L by dense symbolic elimination: after column j,
    rows i, k > j both nonzero in column j make (i, k) nonzero."""
    L = np.tril(P.copy())
    n = L.shape[0]
    for j in range(n):
        rows = np.nonzero(L[j + 1:, j])[0] + j + 1
        if rows.size > 1:
            L[np.ix_(rows, rows)] |= np.tril(np.ones((rows.size, rows.size), bool))
    return L


def etree_from_pattern(L):
    n = L.shape[0]
    par = np.zeros(n, np.int64)
    for j in range(n):
        rows = np.nonzero(L[j + 1:, j])[0]
        par[j] = rows[0] + j + 2 if rows.size else 0
    return par  # 1-based, 0 = root


def check_postordered(par):
    """Parents after children, every subtree a contiguous range ending at its
    root (1-based parent array)."""
    n = len(par)
    size = np.ones(n + 1, np.int64)
    for v in range(1, n + 1):
        p = par[v - 1]
        assert p == 0 or p > v, "parent numbered before its child"
        if p:
            size[p] += size[v]
    for v in range(1, n + 1):
        lo = v - size[v] + 1
        # every node in [lo, v) has an ancestor chain reaching v inside the range
        for u in range(lo, v):
            w = u
            while w and w < v:
                w = par[w - 1]
            assert w == v, "subtree of %d is not contiguous" % v


def dfs_postorder(par, n, order=None):
    """Visiting rule of PostOrderTree (ETree.cpp:59-119) as an explicit DFS:
    roots from the largest label down, children in increasing label order
    (or in the given per-node order); a node is numbered when its subtree is
    done.  Returns invpos (1-based)."""
    kids = [[] for _ in range(n + 1)]
    roots = []
    for v in range(1, n + 1):
        if par[v - 1] == 0 or par[v - 1] == v:
            roots.append(v)
        else:
            kids[par[v - 1]].append(v)
    if order is not None:
        kids = order
    invpos = [0] * n
    counter = 0
    for r in sorted(roots, reverse=True):
        stack = [(r, 0)]
        while stack:
            u, i = stack.pop()
            if i < len(kids[u]):
                stack.append((u, i + 1))
                stack.append((kids[u][i], 0))
            else:
                counter += 1
                invpos[u - 1] = counter
    return invpos


def sorted_children(par, cc):
    """SortChildren's sibling order with explicit lists: children taken in
    decreasing label order go to the front when their count is >= that of
    the current last child, else to the end."""
    n = len(par)
    kids = [[] for _ in range(n + 1)]
    for v in range(n - 1, 0, -1):
        p = par[v - 1]
        if p == 0 or p == v:
            continue
        lst = kids[p]
        if lst and cc[v - 1] < cc[lst[-1] - 1]:
            lst.append(v)
        else:
            lst.insert(0, v)
    return kids


def children_of(par):
    n = len(par)
    kids = [[] for _ in range(n + 1)]
    for v in range(1, n + 1):
        if par[v - 1]:
            kids[par[v - 1]].append(v)
    return kids


def replay_relaxation(L, par, fund, rp):
    """Greedy relaxation (symPACKMatrixBase_impl.hpp:473-609) replayed on the
    dense pattern: fundamental supernodes from the last to the first; one
    joins the relaxed group that starts right after it when its last
    column's parent lies in that group, and the merged width and explicit
    zeros pass the thresholds.  The inputs come from the pattern itself:
    column counts from L, and the zeros a merge introduces counted entry by
    entry (columns of the child padded to the merged structure), which must
    equal the closed form w (cc(parent group's first column) + w - cc(child's
    first column)) of a nested chain.  As in the reference, zeros of merges
    taken because the merged width is within nrelax0 are not added to the
    group's running total.  Returns the relaxed partition (1-based first
    columns, n + 1 last)."""
    n = L.shape[0]
    cc = L.sum(axis=0)
    nf = len(fund) - 1
    width = [fund[k + 1] - fund[k] for k in range(nf)]
    of_col = np.zeros(n + 1, np.int64)
    for k in range(nf):
        of_col[fund[k]:fund[k + 1]] = k
    group_start = list(range(nf))     # representative (first fundamental) of each group
    ncols = list(width)
    zeros = [0] * nf

    def rep(k):
        while group_start[k] != k:
            k = group_start[k]
        return k

    for k in range(nf - 1, -1, -1):
        fc, lc = fund[k], fund[k + 1] - 1
        pc = par[lc - 1]
        if pc == 0:
            continue
        g = rep(int(of_col[pc]))
        if g != k + 1:
            continue
        w, pw = width[k], ncols[g]
        pfc = fund[g]
        W = w + pw
        # zeros the merge introduces, entry by entry: every child column is
        # stored over the merged structure (its rows from the parent group's
        # first column on are the parent's rows)
        prow = np.nonzero(L[pfc - 1:, pfc - 1])[0] + pfc
        merged_rows = set(prow.tolist()) | set(range(fc, pfc))
        newzeros = 0
        for c in range(fc, pfc):
            stored = sum(1 for r in merged_rows if r >= c)
            newzeros += stored - int(cc[c - 1])
        assert newzeros == w * (int(cc[pfc - 1]) + w - int(cc[fc - 1])), "merge zeros are not the chain closed form"
        merge = False
        tot = zeros[g]
        if (W <= rp.maxSize or rp.maxSize == 0) and rp.nrelax0 > 0:
            if W <= rp.nrelax0:
                merge = True
            elif newzeros == 0:
                merge = True
            else:
                # the reference's stored size of the merged supernode, in
                # double arithmetic (it can reach 0 or below when the parent
                # group's first column has fewer entries than the group's width)
                size = np.float64(W * (W + 1) / 2 + W * (int(cc[pfc - 1]) - pw))
                with np.errstate(divide="ignore", invalid="ignore"):
                    z = np.float64(tot + newzeros) / size
                tot += newzeros
                merge = (W <= rp.nrelax1 and z < rp.zrelax0) or (W <= rp.nrelax2 and z < rp.zrelax1) or z < rp.zrelax2
        if merge:
            ncols[k] += ncols[g]
            zeros[k] = tot
            group_start[g] = k
    return [fund[k] for k in range(nf) if rep(k) == k] + [n + 1]


# ------------------------------------------------------------ the matrices
def _matrices():
    out = []
    n, cp, ri, v = M.laplacian_2d_5pt(24)
    out.append(("lap2d_24", n, cp, ri, "MMD"))
    out.append(("lap2d_24_amd", n, cp, ri, "AMD"))
    n, cp, ri, v = M.diffusion_3d_7pt(9, seed=2)
    out.append(("diff3d_9", n, cp, ri, "MMD"))
    out.append(("diff3d_9_nd", n, cp, ri, "NDBOX"))
    n, cp, ri, v = M.stencil_3d_27pt(7)
    out.append(("st27_7", n, cp, ri, "AMD"))
    n, cp, ri, v = M.random_spd(600, 0.006, 5)
    out.append(("rand600", n, cp, ri, "MMD"))
    n, cp, ri, v = M.random_spd(900, 0.003, 8)  # forest-like: many roots and siblings
    out.append(("rand900_sparse", n, cp, ri, "MMD"))
    return out


MATS = _matrices()
RELAX = [dict(), dict(maxsize=12), dict(nrelax=(0, 32, 64)), dict(nrelax=(16, 48, 96), maxsize=40)]


def _run_product(n, cp, ri, ordering, relax):
    rp = sp.RelaxationParameters()
    if "nrelax" in relax:
        rp.nrelax0, rp.nrelax1, rp.nrelax2 = relax["nrelax"]
    if relax.get("maxsize"):
        rp.SetMaxSize(relax["maxsize"])
    A = sp.DistSparseMatrix(n, cp, ri, np.ones(len(ri)))
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering=ordering, relax=rp))
    m.SymbolicFactorization(A)
    perm, invp = m.GetOrdering()
    ccu, ccr = m.GetColumnCounts()
    xl, lx = m.GetStructure()
    xs = list(m.GetSupernodalPartition())
    structs = [list(lx[xl[I] - 1:xl[I + 1] - 1]) for I in range(len(xs) - 1)]
    return dict(perm=list(perm), parent=list(m.GetETree()), cc_unrelaxed=list(ccu), cc=list(ccr),
                xsuper=xs, sup=list(m.GetSupMembership()), structs=structs), rp


def _run_oracle(n, cp, ri, perm0, rp):
    r = dict(nrelax0=rp.nrelax0, nrelax1=rp.nrelax1, nrelax2=rp.nrelax2, max_size=rp.maxSize,
             zrelax0=rp.zrelax0, zrelax1=rp.zrelax1, zrelax2=rp.zrelax2)
    return O.symbolic_chain(n, cp, ri, list(perm0), relax=r)


def _initial_perm(n, cp, ri, ordering):
    invp = np.zeros(n, np.int32)
    perm = np.zeros(n, np.int32)
    lib = sp._lib.load()
    import ctypes
    code = sp._lib.ORDERINGS[ordering]
    cp64 = np.ascontiguousarray(cp, np.int64)
    ri32 = np.ascontiguousarray(ri, np.int32)
    sp._lib.check(lib.spk_compute_ordering(code, n, cp64.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           ri32.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 0,
                                           perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                           invp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "ordering")
    return list(perm)


# --------------------------------------------------------------- the pins
def _pin(res, n, cp, ri, rp, initial_perm):
    perm = res["perm"]
    assert sorted(perm) == list(range(1, n + 1))
    P = dense_pattern(n, cp, ri, perm)
    L = eliminate(P)
    par = np.asarray(res["parent"])
    # etree of the final ordering, and it is postordered
    assert np.array_equal(etree_from_pattern(L), par)
    check_postordered(par)
    # column counts (unrelaxed) from the pattern
    ccu = L.sum(axis=0)
    assert list(ccu) == res["cc_unrelaxed"]

    # postorder + SortChildren relabel one elimination tree: the parent of
    # every vertex is the same vertex under the initial and the final
    # ordering, and each vertex keeps its column count
    P0 = dense_pattern(n, cp, ri, initial_perm)
    L0 = eliminate(P0)
    par0 = etree_from_pattern(L0)
    cc0 = L0.sum(axis=0)
    inv_final = [0] * n
    for i, v in enumerate(perm):
        inv_final[v - 1] = i + 1
    for i0 in range(1, n + 1):
        f = inv_final[initial_perm[i0 - 1] - 1]
        p0 = par0[i0 - 1]
        want = inv_final[initial_perm[p0 - 1] - 1] if p0 else 0
        assert par[f - 1] == want, "not the same elimination tree"
        assert cc0[i0 - 1] == ccu[f - 1]
    # the final ordering is PostOrderTree (ETree.cpp:59-119) then
    # SortChildren (ETree.cpp:122-225), rebuilt here with explicit child
    # lists: the DFS numbers a node after its subtree, roots from the
    # largest label down; PostOrderTree visits children by increasing
    # label; SortChildren inserts children in decreasing label order, in
    # front when their count is >= the current last child's, else at the end
    inv1 = dfs_postorder(par0, n)
    par1 = [0] * n
    for i in range(1, n + 1):
        par1[inv1[i - 1] - 1] = inv1[par0[i - 1] - 1] if par0[i - 1] else 0
    perm1 = [0] * n
    cc1 = [0] * n
    for i in range(1, n + 1):
        perm1[inv1[i - 1] - 1] = initial_perm[i - 1]
        cc1[inv1[i - 1] - 1] = int(cc0[i - 1])
    inv2 = dfs_postorder(np.asarray(par1), n, order=sorted_children(par1, cc1))
    perm2 = [0] * n
    for i in range(1, n + 1):
        perm2[inv2[i - 1] - 1] = perm1[i - 1]
    assert perm2 == perm, "final ordering differs from postorder + SortChildren"
    # SortChildren's effect: the child visited last (numbered right before
    # its parent) has the smallest count among its siblings
    kids = children_of(par)
    for p in range(1, n + 1):
        if len(kids[p]) > 1:
            last = max(kids[p])
            assert last == p - 1
            assert ccu[last - 1] == min(ccu[k - 1] for k in kids[p])

    # fundamental supernodes from the pattern (impl:42-80): column j joins
    # j-1 iff parent(j-1) = j and struct(j-1) minus row j-1 equals struct(j),
    # at most maxSize columns per supernode
    maxs = rp.maxSize
    fund = [1]
    size = 1
    for j in range(2, n + 1):
        same = par[j - 2] == j and np.array_equal(L[j - 1:, j - 2], L[j - 1:, j - 1])
        if same and (maxs == 0 or size < maxs):
            size += 1
        else:
            fund.append(j)
            size = 1
    fund.append(n + 1)
    xs = res["xsuper"]
    # relaxed supernodes are unions of consecutive fundamental supernodes
    assert set(xs) <= set(fund)
    assert xs[0] == 1 and xs[-1] == n + 1
    widths = np.diff(xs)
    if maxs > 0:
        assert widths.max() <= maxs
    # membership array
    for I in range(len(xs) - 1):
        assert all(res["sup"][c - 1] == I + 1 for c in range(xs[I], xs[I + 1]))
    fidx = {f: i for i, f in enumerate(fund)}
    for I in range(len(xs) - 1):
        a, b = fidx[xs[I]], fidx[xs[I + 1]]
        for k in range(a, b - 1):
            # merged child k: the parent of its last column lies in the
            # columns of the fundamental supernodes after it in the union
            last = fund[k + 1] - 1
            pc = par[last - 1]
            assert fund[k + 1] <= pc < xs[I + 1], "relaxed supernode is not a chain union"
    # merges within nrelax0 are always taken (impl:529-533): a boundary
    # between relaxed supernodes R1 | R2 where R1's last fundamental
    # supernode F has its parent column in R2 and |F| + |R2| <= nrelax0
    # (and <= maxSize) would have been merged
    if rp.nrelax0 > 0:
        for I in range(1, len(xs) - 1):
            fk = fund[fidx[xs[I]] - 1]  # first column of the last fundamental of R1
            F = xs[I] - fk
            pc = par[xs[I] - 2]
            R2 = xs[I + 1] - xs[I]
            ms = F + R2
            if xs[I] <= pc < xs[I + 1] and ms <= rp.nrelax0 and (maxs == 0 or ms <= maxs):
                pytest.fail("merge of size %d <= nrelax0 not taken at column %d" % (ms, xs[I]))

    # the relaxation's merge decisions replayed from the pattern
    assert replay_relaxation(L, par, fund, rp) == list(xs), "relaxation decisions differ"

    # structure of every relaxed supernode = union of its columns' patterns
    for I in range(len(xs) - 1):
        fc, lc = xs[I], xs[I + 1] - 1
        rows = np.nonzero(L[fc - 1:, fc - 1:lc].any(axis=1))[0] + fc
        assert list(rows) == res["structs"][I]
        # relaxed column counts follow the structure
        for c in range(fc, lc + 1):
            assert res["cc"][c - 1] == len(rows) - (c - fc)


@pytest.mark.parametrize("relax", RELAX, ids=["default", "max12", "norelax", "relax16_max40"])
@pytest.mark.parametrize("case", MATS, ids=[m[0] for m in MATS])
def test_product_chain_pinned(case, relax):
    name, n, cp, ri, ordering = case
    res, rp = _run_product(n, cp, ri, ordering, relax)
    _pin(res, n, cp, ri, rp, _initial_perm(n, cp, ri, ordering))


@pytest.mark.parametrize("relax", RELAX[:2], ids=["default", "max12"])
@pytest.mark.parametrize("case", MATS[:3] + MATS[5:6], ids=[m[0] for m in MATS[:3] + MATS[5:6]])
def test_oracle_chain_pinned(case, relax):
    name, n, cp, ri, ordering = case
    _, rp = _run_product(n, cp, ri, ordering, relax)
    perm0 = _initial_perm(n, cp, ri, ordering)
    res = _run_oracle(n, cp, ri, perm0, rp)
    _pin(res, n, cp, ri, rp, perm0)


def test_dfs_rule_matches_oracle_postorder():
    """The DFS written here and the oracle's BTreeToPO agree on forests
    with many roots and unsorted sibling labels."""
    rng = np.random.default_rng(4)
    for trial in range(20):
        n = int(rng.integers(5, 60))
        par = [0] * n
        for v in range(1, n + 1):
            if rng.random() < 0.75 and v < n:
                par[v - 1] = int(rng.integers(v + 1, n + 1))
        assert O.postorder(par)[0] == dfs_postorder(np.asarray(par), n)


def test_pins_catch_planted_mistakes():
    """Mistakes planted in oracle copies fail the pins."""
    n, cp, ri, v = M.random_spd(300, 0.01, 9)
    res, rp = _run_product(n, cp, ri, "MMD", {})
    perm0 = _initial_perm(n, cp, ri, "MMD")
    good = _run_oracle(n, cp, ri, perm0, rp)
    _pin(good, n, cp, ri, rp, perm0)
    # (1) a partition that splits a fundamental supernode where none ends
    bad = dict(good)
    xs = list(good["xsuper"])
    w = np.diff(xs)
    I = int(np.argmax(w))
    if w[I] > 1:
        bad["xsuper"] = xs[:I + 1] + [xs[I] + 1] + xs[I + 1:]
        with pytest.raises(AssertionError):
            _pin(bad, n, cp, ri, rp, perm0)
    # (2) a structure missing a row
    bad = dict(good)
    bad["structs"] = [list(s) for s in good["structs"]]
    J = max(range(len(bad["structs"])), key=lambda k: len(bad["structs"][k]))
    bad["structs"][J] = bad["structs"][J][:-1]
    with pytest.raises(AssertionError):
        _pin(bad, n, cp, ri, rp, perm0)
    # (3) siblings in the wrong order (two children of a node swapped)
    par = np.asarray(good["parent"])
    kids = children_of(par)
    p = next(p for p in range(1, n + 1) if len(kids[p]) > 1)
    bad = dict(good)
    bad["cc_unrelaxed"] = list(good["cc_unrelaxed"])
    bad["cc_unrelaxed"][p - 2] += 1
    with pytest.raises(AssertionError):
        _pin(bad, n, cp, ri, rp, perm0)
