"""Oracle: pure-Python restatement of symPACK's symbolic analysis chain.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may use this module, and only as the checker.

Follows, function by function:
  etree_liu          src/ETree.cpp:317-393 (ConstructETree, serial)
  postorder          src/ETree.cpp:25-119  (BTreeToPO, PostOrderTree)
  sort_children      src/ETree.cpp:122-225 (SortChildren)
  column_counts      brute-force symbolic elimination (exact counts; the
                     reference uses Gilbert-Ng-Peyton, impl:309-466, whose
                     result is the same set cardinality)
  find_supernodes    include/sympack/impl/symPACKMatrixBase_impl.hpp:42-80
  relax_supernodes   impl:473-609
  structure          impl:613-938 (row structure = union of children
                     structures beyond their columns and the matrix rows)
  boxnd / gridnd     src/boxnd.f, src/gridnd.f
Pure-Python loops: small problems only (n up to a few thousand).

Parity pins: orderings against the compiled reference Fortran
(oracle/_ref) golden vectors; etree and column counts against brute-force
symbolic elimination (tests/test_symbolic.py); postorder, SortChildren
(the whole final ordering rebuilt with explicit child lists), fundamental
supernodes, the relaxed partition's structure (unions of fundamental
supernodes along parent chains, width caps, every merge of size <= nrelax0
taken) and the relaxed row structures against a dense boolean elimination
(tests/test_symbolic_invariants.py), and relax_supernodes' merge decisions
by replaying them on the dense pattern (each merge's zeros counted entry by
entry, the thresholds applied) in the same test.
"""
import math

import numpy as np


def expand_symmetric(n, colptr, rowind, baseval=0):
    """Full sorted adjacency (incl. diagonal if stored) as python lists, 1-based."""
    adj = [[] for _ in range(n)]
    for c in range(n):
        for p in range(colptr[c] - baseval, colptr[c + 1] - baseval):
            r = int(rowind[p]) - baseval
            adj[c].append(r + 1)
            if r > c:
                adj[r].append(c + 1)
    return [sorted(a) for a in adj]


def compose(perm, invp, invp2):
    """Ordering::Compose (src/Ordering.cpp:1301-1315), 1-based lists."""
    n = len(invp)
    invp = [invp2[invp[i] - 1] for i in range(n)]
    perm = [0] * n
    for i in range(n):
        perm[invp[i] - 1] = i + 1
    return perm, invp


def etree_liu(n, adj, perm, invp):
    """Elimination tree of P A P^T (1-based parents, 0 = root)."""
    parent = [0] * n
    ancstr = [0] * n
    for i in range(1, n + 1):
        node = perm[i - 1]
        col = adj[node - 1]
        if len(col) < 2:
            continue
        for r in col:
            nbr = invp[r - 1]
            if nbr >= i:
                continue
            if ancstr[nbr - 1] == i:
                continue
            brk = False
            while ancstr[nbr - 1] > 0:
                if ancstr[nbr - 1] == i:
                    brk = True
                    break
                nxt = ancstr[nbr - 1]
                ancstr[nbr - 1] = i
                nbr = nxt
            if not brk:
                parent[nbr - 1] = i
                ancstr[nbr - 1] = i
    return parent


def _btree_to_po(n, fson, brother):
    stack = []
    invpos = [0] * n
    vertex, m = n, 0
    while m < n:
        while vertex > 0:
            stack.append(vertex)
            vertex = fson[vertex - 1]
        done = False
        while vertex == 0:
            if not stack:
                done = True
                break
            vertex = stack.pop()
            m += 1
            invpos[vertex - 1] = m
            vertex = brother[vertex - 1]
        if done:
            break
    return invpos


def postorder(parent):
    """PostOrderTree: returns (invpos, postordered parent)."""
    n = len(parent)
    fson = [0] * n
    brother = [0] * n
    lroot = n
    for vertex in range(n - 1, 0, -1):
        cur = parent[vertex - 1]
        if cur == 0 or cur == vertex:
            brother[lroot - 1] = vertex
            lroot = vertex
        else:
            brother[vertex - 1] = fson[cur - 1]
            fson[cur - 1] = vertex
    invpos = _btree_to_po(n, fson, brother)
    po = [0] * n
    for i in range(1, n + 1):
        nd = parent[i - 1]
        po[invpos[i - 1] - 1] = invpos[nd - 1] if nd > 0 else 0
    return invpos, po


def sort_children(poparent, cc):
    """SortChildren: returns (invpos, new parent, permuted cc)."""
    n = len(poparent)
    fson = [0] * n
    brother = [0] * n
    lson = [0] * n
    lroot = n
    for vertex in range(n - 1, 0, -1):
        cur = poparent[vertex - 1]
        if cur == 0 or cur == vertex:
            brother[lroot - 1] = vertex
            lroot = vertex
        else:
            nd = lson[cur - 1]
            if nd > 0:
                if cc[vertex - 1] >= cc[nd - 1]:
                    brother[vertex - 1] = fson[cur - 1]
                    fson[cur - 1] = vertex
                else:
                    brother[nd - 1] = vertex
                    lson[cur - 1] = vertex
            else:
                fson[cur - 1] = vertex
                lson[cur - 1] = vertex
    brother[lroot - 1] = 0
    invpos = _btree_to_po(n, fson, brother)
    newpar = [0] * n
    for i in range(1, n + 1):
        nd = poparent[i - 1]
        newpar[invpos[i - 1] - 1] = invpos[nd - 1] if nd > 0 else 0
    newcc = [0] * n
    for i in range(n):
        newcc[invpos[i] - 1] = cc[i]
    return invpos, newpar, newcc


def column_counts(n, adj, perm, invp):
    """Exact column counts of L for P A P^T by symbolic elimination, and the
    etree parent from the same elimination (brute force, small n)."""
    struct = [set() for _ in range(n)]
    for j in range(n):
        for r in adj[perm[j] - 1]:
            pr = invp[r - 1] - 1
            if pr > j:
                struct[j].add(pr)
    cc = [0] * n
    parent = [0] * n
    for j in range(n):
        rows = sorted(struct[j])
        cc[j] = len(rows) + 1
        if rows:
            p = rows[0]
            parent[j] = p + 1
            struct[p].update(rows[1:])
    return cc, parent


def find_supernodes(poparent, cc, max_size=0):
    n = len(cc)
    sup = [0] * n
    nsuper, size = 1, 1
    sup[0] = 1
    for i in range(2, n + 1):
        if poparent[i - 2] == i and cc[i - 2] == cc[i - 1] + 1 and (size < max_size or max_size == 0):
            size += 1
            sup[i - 1] = nsuper
            continue
        nsuper += 1
        size = 1
        sup[i - 1] = nsuper
    xsuper = [1] + [0] * nsuper
    last = nsuper + 1
    for i in range(n, 0, -1):
        k = sup[i - 1]
        if k != last:
            xsuper[last - 1] = i + 1
        last = k
    xsuper[0] = 1
    return sup, xsuper


def relax_supernodes(poparent, cc, sup, xsuper, nrelax0=8, nrelax1=32, nrelax2=64, max_size=0,
                     zrelax0=0.8, zrelax1=0.1, zrelax2=0.05):
    """Returns (cc, sup, xsuper) after relaxation."""
    nsuper = len(xsuper) - 1
    root = list(range(1, nsuper + 1))     # set root label (lowest member)
    rep = list(range(1, nsuper + 1))      # union-find parent
    ncols = [xsuper[k] - xsuper[k - 1] for k in range(1, nsuper + 1)]
    zeros = [0] * nsuper
    newcc = [cc[xsuper[k - 1] - 1] for k in range(1, nsuper + 1)]

    def find(i):
        while rep[i - 1] != i:
            i = rep[i - 1]
        return i

    for ksup in range(nsuper, 0, -1):
        fst, lst = xsuper[ksup - 1], xsuper[ksup] - 1
        width = ncols[ksup - 1]
        pcol = poparent[lst - 1]
        if pcol == 0:
            continue
        psn = root[find(sup[pcol - 1]) - 1]
        if psn != ksup + 1:
            continue
        pw = ncols[psn - 1]
        pfc = xsuper[psn - 1]
        totzeros = zeros[psn - 1]
        msize = width + pw
        merge = False
        if (msize <= max_size or max_size == 0) and nrelax0 > 0:
            if msize <= nrelax0:
                merge = True
            elif msize <= max_size or max_size == 0:
                nnzchild = float(cc[fst - 1])
                nnzparent = float(cc[pfc - 1])
                xnew = width * (nnzparent + width - nnzchild)
                if xnew == 0:
                    merge = True
                else:
                    xtot = float(totzeros) + xnew
                    xm = float(msize)
                    xsize = xm * (xm + 1) / 2 + xm * (nnzparent - pw)
                    # IEEE semantics of the reference's double division
                    if xsize != 0.0:
                        z = xtot / xsize
                    elif xtot == 0.0:
                        z = float("nan")
                    else:
                        z = math.copysign(float("inf"), xtot)
                    totzeros += int(xnew)
                    if msize <= nrelax1 and z < zrelax0:
                        merge = True
                    elif msize <= nrelax2 and z < zrelax1:
                        merge = True
                    elif z < zrelax2:
                        merge = True
        if merge:
            ncols[ksup - 1] += ncols[psn - 1]
            zeros[ksup - 1] = totzeros
            newcc[ksup - 1] = width + newcc[psn - 1]
            t = find(psn)
            rep[ksup - 1] = t
            root[t - 1] = ksup
    newx, ncc = [], []
    for ksup in range(1, nsuper + 1):
        if root[find(ksup) - 1] == ksup:
            newx.append(xsuper[ksup - 1])
            ncc.append(newcc[ksup - 1])
    newx.append(xsuper[nsuper])
    cc = list(cc)
    sup = list(sup)
    for k in range(1, len(newx)):
        for col in range(newx[k - 1], newx[k]):
            sup[col - 1] = k
            cc[col - 1] = ncc[k - 1] - (col - newx[k - 1])
    return cc, sup, newx


def symbolic_chain(n, colptr, rowind, perm0, baseval=0, relax=None, ordering_kind="sgraph"):
    """Full chain from an initial ordering (perm0: perm[new-1] = old, 1-based
    list) to the relaxed supernodes and structures.  Returns a dict."""
    adj = expand_symmetric(n, colptr, rowind, baseval)
    perm = list(perm0)
    invp = [0] * n
    for i, p in enumerate(perm):
        invp[p - 1] = i + 1
    parent = etree_liu(n, adj, perm, invp)
    invpos, po = postorder(parent)
    perm, invp = compose(perm, invp, invpos)
    cc, _ = column_counts(n, adj, perm, invp)
    if ordering_kind != "natural":
        invpos2, po, cc = sort_children(po, cc)
        perm, invp = compose(perm, invp, invpos2)
    cc_unrelaxed = list(cc)
    sup, xsuper = find_supernodes(po, cc, (relax or {}).get("max_size", 0))
    r = dict(nrelax0=8, nrelax1=32, nrelax2=64, max_size=0, zrelax0=0.8, zrelax1=0.1, zrelax2=0.05)
    if relax:
        r.update(relax)
    if r["nrelax0"] > 0:
        cc, sup, xsuper = relax_supernodes(po, cc, sup, xsuper, **r)
    # structures: union of children (beyond their columns) + own columns +
    # permuted matrix rows below the first column
    nsuper = len(xsuper) - 1
    structs = []
    kids = [[] for _ in range(nsuper)]
    for k in range(1, nsuper + 1):
        fst, lst = xsuper[k - 1], xsuper[k] - 1
        rows = set(range(fst, lst + 1))
        for j in kids[k - 1]:
            jw = xsuper[j] - xsuper[j - 1]
            rows.update(structs[j - 1][jw:])
        for col in range(fst, lst + 1):
            for r0 in adj[perm[col - 1] - 1]:
                pr = invp[r0 - 1]
                if pr > fst:
                    rows.add(pr)
        s = sorted(rows)
        structs.append(s)
        if len(s) > lst - fst + 1:
            pcol = po[lst - 1]
            kids[sup[pcol - 1] - 1].append(k)
    flops = float(sum(float(c) ** 2 for c in cc_unrelaxed))
    return dict(perm=perm, invp=invp, parent=po, cc_unrelaxed=cc_unrelaxed, cc=cc, sup=sup,
                xsuper=xsuper, structs=structs, flops=flops, nnzL=sum(cc_unrelaxed))


def boxnd(p, q, r):
    """boxnd.f: inverse ordering of a p x q x r box (x fastest), 1-based."""
    box = np.zeros(p * q * r, np.int64)

    def at(ii, jj, kk):
        return (ii - 1) + p * (jj - 1) + p * q * (kk - 1)

    last = p * q * r
    stack = [(1, 1, 1, p, q, r)]
    while stack:
        i, j, k, m, nn, l = stack.pop()
        if m >= nn and m >= l:
            dim = 0
        elif nn >= m and nn >= l:
            dim = 1
        else:
            dim = 2
        size = (m, nn, l)[dim]
        if size <= 2:
            for ii in range(i + m - 1, i - 1, -1):
                for jj in range(j + nn - 1, j - 1, -1):
                    for kk in range(k + l - 1, k - 1, -1):
                        box[at(ii, jj, kk)] = last
                        last -= 1
            continue
        half = (size - 1) // 2
        if dim == 0:
            mid = i + half
            stack.append((i, j, k, half, nn, l))
            stack.append((mid + 1, j, k, m - half - 1, nn, l))
            for jj in range(j + nn - 1, j - 1, -1):
                for kk in range(k + l - 1, k - 1, -1):
                    box[at(mid, jj, kk)] = last
                    last -= 1
        elif dim == 1:
            mid = j + half
            stack.append((i, j, k, m, half, l))
            stack.append((i, mid + 1, k, m, nn - half - 1, l))
            for ii in range(i + m - 1, i - 1, -1):
                for kk in range(k + l - 1, k - 1, -1):
                    box[at(ii, mid, kk)] = last
                    last -= 1
        else:
            mid = k + half
            stack.append((i, j, k, m, nn, half))
            stack.append((i, j, mid + 1, m, nn, l - half - 1))
            for ii in range(i + m - 1, i - 1, -1):
                for jj in range(j + nn - 1, j - 1, -1):
                    box[at(ii, jj, mid)] = last
                    last -= 1
    return box


def gridnd(p, q):
    """gridnd.f: inverse ordering of a p x q grid, 1-based."""
    grid = np.zeros(p * q, np.int64)
    last = p * q
    stack = [(1, 1, p, q)]
    while stack:
        i, j, m, nn = stack.pop()
        if m >= nn:
            if m <= 2:
                for ii in range(i + m - 1, i - 1, -1):
                    for jj in range(j + nn - 1, j - 1, -1):
                        grid[(ii - 1) + p * (jj - 1)] = last
                        last -= 1
            else:
                half = (m - 1) // 2
                mid = i + half
                stack.append((i, j, half, nn))
                stack.append((mid + 1, j, m - half - 1, nn))
                for jj in range(j + nn - 1, j - 1, -1):
                    grid[(mid - 1) + p * (jj - 1)] = last
                    last -= 1
        else:
            if nn <= 2:
                for ii in range(i + m - 1, i - 1, -1):
                    for jj in range(j + nn - 1, j - 1, -1):
                        grid[(ii - 1) + p * (jj - 1)] = last
                        last -= 1
            else:
                half = (nn - 1) // 2
                mid = j + half
                stack.append((i, j, m, half))
                stack.append((i, mid + 1, m, nn - half - 1))
                for ii in range(i + m - 1, i - 1, -1):
                    grid[(ii - 1) + p * (mid - 1)] = last
                    last -= 1
    return grid


def invp_to_perm(invp):
    n = len(invp)
    perm = [0] * n
    for i in range(n):
        perm[int(invp[i]) - 1] = i + 1
    return perm


def nd_box_perm(n):
    k = int(math.ceil(math.pow(n, 1.0 / 3.0)))
    return invp_to_perm(boxnd(k, k, k))
