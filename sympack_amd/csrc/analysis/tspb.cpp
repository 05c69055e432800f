// Supernode reordering by the traveling-salesman heuristic ("-refine TSPB",
// symPACKMatrix2D.hpp:3113-3263): inside every supernode the columns are
// put in the order of a tour that visits them with few changes of row
// subtree, so that the off-diagonal structure of the supernodes below falls
// into fewer runs of consecutive rows.  The reference calls Ng and Peyton's
// Fortran (src/TSPB/ordsup_ind_tsp_paths[2].f and the routines below it);
// the steps here are that method's, one function per routine, with the same
// visiting orders and tie-breaking, because the resulting ordering fixes the
// symbolic structure and must be reproduced exactly:
//  * skeleton graph of the factor (fnskel.f): for every column, the
//    neighbours of lower number that are leaves of its row subtree (back
//    list) and, per column, the higher columns it is such a leaf for
//    (forward list, ascending);
//  * optional compression (compress_skel2.f, code TSPB): columns of one
//    supernode with identical back lists are represented by the first of
//    them, found by hashing the list sums;
//  * per supernode R (bottom up): the postordered subtree of the supernodes
//    whose structure reaches R (post_suptree.f, sup_etpost.f), the row-count
//    distances between R's columns over that subtree (ordsup_tsp_paths.f:
//    least-common-ancestor path counting on a disjoint-set forest with the
//    reference's path-halving finds), and the nearest-insertion tour through
//    R's columns starting from a virtual column 0 (near_ins[2].f);
//  * the tour order numbers R's columns; the new positions are composed into
//    the ordering by the caller (invinv.f).
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "analysis.hpp"

namespace spk {

namespace {

// fnskel.f.  Lists live in one array with the graph's column pointers: the
// back list of old column c fills its segment from the end, the forward
// list from the start (their lengths add up to at most the degree of c).
struct Skeleton {
  std::vector<Ptr> x;       // segment starts, 1-based, n+1
  std::vector<Int> adj;     // old column numbers
  std::vector<Int> lenf;    // forward list length per old column (1-based index)
  std::vector<Int> lenb;    // back list length per old column
};

void build_skeleton(const Graph& g, const std::vector<Int>& perm, const std::vector<Int>& invp,
                    const std::vector<Int>& etpar, Skeleton& s) {
  const Int n = g.n;
  s.x.assign(g.colptr.begin(), g.colptr.end());
  s.adj.assign(g.rowind.size() + 1, 0);
  s.lenf.assign(n + 1, 0);
  s.lenb.assign(n + 1, 0);
  std::vector<Int> fdesc(n + 1), prvnbr(n + 1, 0);
  for (Int k = 0; k <= n; ++k) fdesc[k] = k;
  // first descendant of every column in the (postordered) elimination tree
  for (Int k = 1; k <= n; ++k) {
    const Int p = etpar[k];
    if (fdesc[k] < fdesc[p]) fdesc[p] = fdesc[k];
  }
  auto at = [&](Ptr pos) -> Int& { return s.adj[pos]; };  // 1-based position
  for (Int low = 1; low <= n; ++low) {
    const Int first = fdesc[low];
    const Int oldlow = perm[low];
    for (Ptr q = g.colptr[oldlow - 1]; q < g.colptr[oldlow]; ++q) {
      const Int oldhi = g.rowind[q - 1];
      const Int hi = invp[oldhi];
      if (hi <= low) continue;
      // low starts a new branch of hi's row subtree when hi's previous
      // neighbour is not in low's subtree: low is a leaf of it
      if (first > prvnbr[hi]) {
        ++s.lenb[oldhi];
        at(s.x[oldhi] - s.lenb[oldhi]) = oldlow;
      }
      prvnbr[hi] = low;
    }
  }
  // forward lists, ascending in the new numbering of their entries
  for (Int hi = 1; hi <= n; ++hi) {
    const Int oldhi = perm[hi];
    for (Ptr j = s.x[oldhi] - 1; j >= s.x[oldhi] - s.lenb[oldhi]; --j) {
      const Int oldlow = at(j);
      ++s.lenf[oldlow];
      at(s.x[oldlow - 1] + s.lenf[oldlow] - 1) = oldhi;
    }
  }
}

// compress_skel2.f: rep[new] = representative column (new numbering).
void compress_skeleton(Int n, const std::vector<Int>& snode, const std::vector<Int>& invp, Skeleton& s,
                       std::vector<Int>& rep) {
  rep.resize(n + 1);
  std::vector<Int> heads(n + 1, 0), next(n + 1, 0), stamp(n + 1, 0);
  for (Int i = 1; i <= n; ++i) rep[i] = i;
  auto back_begin = [&](Int c) { return s.x[c] - 1; };            // last position, walked downwards
  auto back_end = [&](Int c) { return s.x[c] - s.lenb[c]; };       // first position
  for (Int c = 1; c <= n; ++c) {
    if (s.lenb[c] == 0) continue;
    const Int newc = invp[c];
    int64_t sum = 0;
    for (Ptr k = back_begin(c); k >= back_end(c); --k) {
      const Int v = s.adj[k];
      sum += v;
      stamp[v] = c;
    }
    const Int bin = (Int)(sum % (int64_t)n) + 1;
    bool absorbed = false;
    for (Int d = heads[bin]; d > 0; d = next[d]) {
      const Int newd = invp[d];
      if (snode[newd] != snode[newc] || s.lenb[d] != s.lenb[c]) continue;
      bool same = true;
      for (Ptr k = back_begin(d); k >= back_end(d); --k)
        if (stamp[s.adj[k]] != c) same = false;
      if (same) {
        rep[newc] = newd;
        absorbed = true;
        break;
      }
    }
    if (absorbed) continue;
    next[c] = heads[bin];
    heads[bin] = c;
  }
  // forward lists keep representatives only
  for (Int c = 1; c <= n; ++c) {
    if (s.lenf[c] == 0) continue;
    Ptr w = s.x[c - 1];
    for (Ptr k = s.x[c - 1]; k < s.x[c - 1] + s.lenf[c]; ++k) {
      const Int v = s.adj[k];
      const Int nv = invp[v];
      if (rep[nv] == nv) s.adj[w++] = v;
    }
    s.lenf[c] = (Int)(w - s.x[c - 1]);
  }
}

struct Structure {
  const std::vector<Int>& xsuper;   // 1-based, index 1..nsuper+1
  const std::vector<Ptr>& xlindx;
  const std::vector<Idx>& lindx;
  const std::vector<Int>& snode;    // 1-based columns
};

// post_suptree.f + sup_etpost.f: the supernodes linked to root (and root),
// postordered into sperm[1..ns2] with the parent vector suppar in the new
// labels; advances fstloc of every linked supernode past its rows in root.
Int postorder_subtree(Int root, const Structure& st, const std::vector<Int>& link, std::vector<Ptr>& fstloc,
                      std::vector<Int>& sperm, std::vector<Int>& suppar, std::vector<Int>& parent,
                      std::vector<Int>& fchild, std::vector<Int>& sibling, std::vector<Int>& stack,
                      std::vector<Int>& sinvp) {
  parent[root] = 0;
  fchild[root] = 0;
  sibling[root] = 0;
  Int ns2 = 1;
  for (Int s = link[root]; s > 0; s = link[s]) {
    ++ns2;
    fchild[s] = 0;
    sibling[s] = 0;
    const Ptr first = st.xlindx[s] + (st.xsuper[s + 1] - st.xsuper[s]);
    if (first <= st.xlindx[s + 1] - 1) parent[s] = st.snode[st.lindx[first]];
  }
  for (Int s = link[root]; s > 0; s = link[s]) {
    Ptr i = fstloc[s];
    const Ptr last = st.xlindx[s + 1] - 1;
    while (i <= last && st.snode[st.lindx[i]] == root) ++i;
    fstloc[s] = i;
    const Int p = parent[s];
    sibling[s] = fchild[p];
    fchild[p] = s;
  }
  // postorder by first-son / brother traversal with an explicit stack
  Int count = 0, top = 0, node = root;
  for (;;) {
    while (node > 0) {
      stack[++top] = node;
      node = fchild[node];
    }
    bool descend = false;
    while (top > 0) {
      node = stack[top--];
      sperm[++count] = node;
      sinvp[node] = count;
      node = sibling[node];
      if (node > 0) {
        descend = true;
        break;
      }
    }
    if (!descend) break;
  }
  for (Int k = 1; k <= ns2; ++k) {
    const Int p = parent[sperm[k]];
    suppar[k] = p == 0 ? 0 : sinvp[p];
  }
  return ns2;
}

// ordsup_tsp_paths.f: row-count distances between the nnodes columns of
// root (local labels 1..nnodes, 0 = virtual start) over the postordered
// subtree sperm/suppar.  dist is (nnodes+1)^2, column-major; during the
// sweep its strict lower triangle counts the supernodes common to the row
// subtrees of two columns and its strict upper triangle holds the previous
// leaf of each pair's intersection subtree (one array, as in the reference).
void tsp_distances(Int nnodes, Int root, Int ns2, const Structure& st, const Skeleton& sk,
                   const std::vector<Int>& perm, const std::vector<Int>& invp, const std::vector<Int>& sperm,
                   const std::vector<Int>& suppar, std::vector<Ptr>& fstloc2, std::vector<int32_t>& dist) {
  const int64_t ld = nnodes + 1;
  auto D = [&](Int i, Int j) -> int32_t& { return dist[(size_t)(i + j * ld)]; };
  std::vector<Int> level(ns2 + 1), fdesc(ns2 + 1), set(ns2 + 1);
  level[0] = 0;
  for (Int k = ns2; k >= 1; --k) {
    fdesc[k] = k;
    set[k] = k;
    level[k] = level[suppar[k]] + 1;
  }
  const Int i0 = st.xsuper[root] - 1;  // local label = new column - i0
  std::vector<Int> rowcnt(nnodes + 1, 1), prvlf(nnodes + 1, 0), curlf(nnodes + 1, 0);
  for (Int i = 1; i <= nnodes; ++i) {
    for (Int j = 1; j < i; ++j) D(i, j) = 1;
    for (Int j = i + 1; j <= nnodes; ++j) D(i, j) = 0;
  }
  fdesc[0] = 0;
  for (Int k = 1; k <= ns2; ++k) {
    const Int p = suppar[k];
    if (fdesc[k] < fdesc[p]) fdesc[p] = fdesc[k];
  }
  // find with the reference's halving: set[] is rewritten along the way
  auto find = [&](Int x) {
    Int a = x, b = set[a], c = set[b];
    while (c != b) {
      set[a] = c;
      a = c;
      b = set[a];
      c = set[b];
    }
    return c;
  };
  // "visited" list of columns, in order of their latest visit
  std::vector<Int> lf(nnodes + 1, -1), lb(nnodes + 1, -1);
  Int lhead = 0, ltail = 0;
  for (Int k = 1; k <= ns2; ++k) {
    const Int s = sperm[k];
    const Int oldfst = perm[st.xsuper[s]];
    const Ptr floc = fstloc2[oldfst];
    Ptr lloc = sk.x[oldfst - 1] + sk.lenf[oldfst] - 1;
    // the columns of root among the forward skeleton neighbours of s's first
    // column: a prefix of the remaining list (it is ascending)
    bool any = false;
    if (floc <= lloc) {
      Ptr p = floc;
      for (; p <= lloc; ++p) {
        const Int nw = invp[sk.adj[p]];
        if (st.snode[nw] != root) break;
        any = true;
        curlf[nw - i0] = k;
      }
      lloc = p - 1;
    }
    if (any) {
      fstloc2[oldfst] = lloc + 1;
      for (Ptr p = floc; p <= lloc; ++p) {
        const Int i = invp[sk.adj[p]] - i0;
        const Int pleaf = prvlf[i];
        if (pleaf == 0)
          rowcnt[i] += level[k] - level[ns2];
        else
          rowcnt[i] += level[k] - level[find(pleaf)];
        const Int right = lf[i], left = lb[i];
        for (Int j = (left == -1) ? lhead : right; j > 0; j = lf[j]) {
          const Int nleaf = (fdesc[k] <= curlf[j]) ? k : find(prvlf[j]);
          const Int pl = (j < i) ? D(j, i) : D(i, j);
          if (nleaf == pl) continue;
          Int anc;
          if (pl == 0)
            anc = ns2;
          else if (nleaf < pl)
            anc = pl;
          else
            anc = find(pl);
          if (j < i) {
            D(i, j) += level[nleaf] - level[anc];
            D(j, i) = nleaf;
          } else {
            D(j, i) += level[nleaf] - level[anc];
            D(i, j) = nleaf;
          }
        }
        // move i to the tail of the visited list
        if (left != -1) {
          if (left > 0) lf[left] = right;
          else lhead = right;
          if (right > 0) lb[right] = left;
          else ltail = left;
        }
        if (ltail > 0) {
          lf[ltail] = i;
          lb[i] = ltail;
          lf[i] = 0;
          ltail = i;
        } else {
          lhead = ltail = i;
          lf[i] = 0;
          lb[i] = 0;
        }
      }
      for (Ptr p = floc; p <= lloc; ++p) prvlf[invp[sk.adj[p]] - i0] = k;
    }
    set[k] = suppar[k];
  }
  for (Int i = 1; i <= nnodes; ++i) {
    D(i, 0) = rowcnt[i];
    D(0, i) = rowcnt[i];
    for (Int j = 1; j < i; ++j) {
      D(i, j) = rowcnt[i] + rowcnt[j] - 2 * D(i, j);
      D(j, i) = D(i, j);
    }
  }
}

// near_ins.f / near_ins2.f: nearest-insertion tour from virtual column 0
// through root's columns (representatives only when rep is given; the
// others follow their representative, the later ones nearer to it), then
// the columns are numbered in tour order from colnum + 1.
void nearest_insertion(Int nnodes, Int root, Int ns2, const Structure& st, const std::vector<int32_t>& dist,
                       const std::vector<Int>* rep, Int& colnum, std::vector<Int>& invp2) {
  const int64_t ld = nnodes + 1;
  auto D = [&](Int i, Int j) { return dist[(size_t)(i + j * ld)]; };
  const Int i0 = st.xsuper[root] - 1;
  auto is_rep = [&](Int i) { return rep == nullptr || (*rep)[i0 + i] == i0 + i; };
  std::vector<Int> mdhead(ns2 + 1, 0), mdf(nnodes + 1, 0), mdb(nnodes + 1, 0), md(nnodes + 1, 0);
  std::vector<Int> of(nnodes + 1, 0), ob(nnodes + 1, 0);
  std::vector<Int> lf(nnodes + 1, 0), lb(nnodes + 1, 0);
  // columns not yet in the tour
  Int outhead = 0, todo = 0;
  if (rep == nullptr) {
    for (Int i = nnodes - 1; i >= 1; --i) {
      of[i] = i + 1;
      ob[i + 1] = i;
    }
    of[nnodes] = 0;
    ob[1] = 0;
    outhead = 1;
    todo = nnodes;
  } else {
    for (Int i = 1; i <= nnodes; ++i) {
      if (!is_rep(i)) continue;
      of[i] = outhead;
      if (outhead > 0) ob[outhead] = i;
      outhead = i;
      ob[i] = 0;
      ++todo;
    }
  }
  const Int total = todo;
  const Int lhead = 0;
  lb[0] = -1;
  lf[0] = -1;
  Int mindist = ns2;
  for (Int i = 1; i <= nnodes; ++i) {
    if (!is_rep(i)) continue;
    const Int d = D(i, 0);
    md[i] = d;
    if (d < mindist) mindist = d;
    const Int nx = mdhead[d];
    mdhead[d] = i;
    mdf[i] = nx;
    if (nx > 0) mdb[nx] = i;
    mdb[i] = 0;
  }
  while (todo >= 1) {
    while (mdhead[mindist] == 0) ++mindist;
    const Int i = mdhead[mindist];
    Int nx = mdf[i];
    mdhead[mindist] = nx;
    if (nx > 0) mdb[nx] = 0;
    Int next = of[i], prev = ob[i];
    if (next > 0) ob[next] = prev;
    if (prev > 0) of[prev] = next;
    else outhead = next;
    if (todo == total) {
      // second column of the tour: after 0
      lf[lhead] = i;
      lb[i] = lhead;
      lf[i] = -1;
    } else {
      Int mincost = 2 * ns2;
      for (Int j = lhead; j >= 0;) {
        Int nj = lf[j];
        if (nj == -1) nj = lhead;
        const Int cost = D(i, j) + D(i, nj) - D(j, nj);
        if (cost < mincost) {
          mincost = cost;
          prev = j;
          next = nj;
        }
        j = (nj == lhead) ? -1 : nj;
      }
      lf[prev] = i;
      lb[i] = prev;
      if (next == lhead) {
        lf[i] = -1;
      } else {
        lf[i] = next;
        lb[next] = i;
      }
    }
    // distances of the columns still out of the tour
    for (Int j = outhead; j > 0; j = of[j]) {
      const Int d = D(i, j);
      if (d >= md[j]) continue;
      const Int old = md[j];
      const Int pj = mdb[j], nj = mdf[j];
      if (pj == 0) mdhead[old] = nj;
      else mdf[pj] = nj;
      if (nj > 0) mdb[nj] = pj;
      const Int h = mdhead[d];
      mdf[j] = h;
      if (h > 0) mdb[h] = j;
      mdb[j] = 0;
      mdhead[d] = j;
      md[j] = d;
      if (d < mindist) mindist = d;
    }
    --todo;
  }
  if (rep != nullptr) {
    for (Int i = 1; i <= nnodes; ++i) {
      const Int r = (*rep)[i0 + i];
      if (r == i0 + i) continue;
      const Int rl = r - i0;
      const Int nx = lf[rl];
      lf[rl] = i;
      lb[i] = rl;
      if (nx > 0) lb[nx] = i;
      lf[i] = nx;
    }
  }
  for (Int j = lf[0]; j >= 0; j = lf[j]) invp2[i0 + j] = ++colnum;
}

}  // namespace

std::vector<Int> refine_supernodes_tsp(bool compress, Int n, const Graph& graph, const ETree& et,
                                       const Ordering& ord, const std::vector<Int>& xsuper0,
                                       const std::vector<Int>& snode0, const std::vector<Ptr>& xlindx0,
                                       const std::vector<Idx>& lindx0) {
  const Int nsuper = (Int)xsuper0.size() - 1;
  Int supsiz = 0;
  for (Int s = 0; s < nsuper; ++s) supsiz = std::max(supsiz, xsuper0[s + 1] - xsuper0[s]);
  // the reference holds the distances of the widest supernode in one
  // (supsiz+1)^2 array; refuse what it could not allocate either
  if ((int64_t)(supsiz + 1) * (supsiz + 1) > ((int64_t)1 << 31))
    throw std::logic_error("TSPB supernode reordering: supernode of " + std::to_string(supsiz) +
                           " columns is too wide for its (width+1)^2 distance table");
  // 1-based views (index 0 unused) of the structure and the ordering
  std::vector<Int> xsuper(nsuper + 2), snode(n + 1), perm(n + 1), invp(n + 1), etpar(n + 1);
  std::vector<Ptr> xlindx(nsuper + 2);
  std::vector<Idx> lindx(lindx0.size() + 1);
  for (Int s = 0; s <= nsuper; ++s) {
    xsuper[s + 1] = xsuper0[s];
    xlindx[s + 1] = xlindx0[s];
  }
  for (size_t q = 0; q < lindx0.size(); ++q) lindx[q + 1] = lindx0[q];
  for (Int i = 0; i < n; ++i) {
    snode[i + 1] = snode0[i];
    perm[i + 1] = ord.perm[i];
    invp[i + 1] = ord.invp[i];
    etpar[i + 1] = et.PostParent(i);
  }
  Graph g = graph;
  set_keep_diag(g, false);
  Skeleton sk;
  build_skeleton(g, perm, invp, etpar, sk);
  std::vector<Int> rep;
  if (compress) compress_skeleton(n, snode, invp, sk, rep);

  const Structure st{xsuper, xlindx, lindx, snode};
  std::vector<Ptr> fstloc2(n + 1), fstloc(nsuper + 1, 0);
  for (Int c = 1; c <= n; ++c) fstloc2[c] = sk.x[c - 1];
  std::vector<Int> link(nsuper + 1, 0), sperm(nsuper + 1), suppar(nsuper + 1), parent(nsuper + 1, 0),
      fchild(nsuper + 1, 0), sibling(nsuper + 1, 0), stack(nsuper + 1, 0), sinvp(nsuper + 1, 0);
  std::vector<Int> invp2(n + 1, 0);
  std::vector<int32_t> dist;
  Int colnum = 0;
  for (Int root = 1; root <= nsuper; ++root) {
    const Int ns2 = postorder_subtree(root, st, link, fstloc, sperm, suppar, parent, fchild, sibling, stack,
                                      sinvp);
    const Int nnodes = xsuper[root + 1] - xsuper[root];
    dist.assign((size_t)(nnodes + 1) * (nnodes + 1), 0);
    tsp_distances(nnodes, root, ns2, st, sk, perm, invp, sperm, suppar, fstloc2, dist);
    nearest_insertion(nnodes, root, ns2, st, dist, compress ? &rep : nullptr, colnum, invp2);
    // link the subtree's supernodes to the next root their rows reach
    for (Int k = 1; k <= ns2 - 1; ++k) {
      const Int s = sperm[k];
      if (fstloc[s] < xlindx[s + 1]) {
        const Int nxt = snode[lindx[fstloc[s]]];
        link[s] = link[nxt];
        link[nxt] = s;
      }
    }
    const Ptr first = xlindx[root] + nnodes;
    if (first <= xlindx[root + 1] - 1) {
      const Int nxt = snode[lindx[first]];
      link[root] = link[nxt];
      link[nxt] = root;
      fstloc[root] = first;
    }
  }
  return std::vector<Int>(invp2.begin() + 1, invp2.end());
}

}  // namespace spk
