// Graph utilities and elimination tree, restating src/DistSparseMatrixGraph.cpp
// (ExpandSymmetric, SetKeepDiag, Permute on one rank) and src/ETree.cpp.
#include <algorithm>
#include <numeric>

#include "analysis.hpp"

namespace spk {

// DistSparseMatrixGraph::ExpandSymmetric (cpp:929-1083) on one rank.  Every
// strictly lower entry (row>col) is mirrored; mirrored entries are placed in
// front of the column's own entries and the column is then sorted (the
// reference sorts when the input is unsorted; the driver input is sorted).
Graph expand_symmetric(Int n, const Ptr* colptr, const Idx* rowind, int baseval) {
  Graph g;
  g.n = n;
  std::vector<Ptr> cnt(n + 1, 0);
  for (Int c = 0; c < n; ++c) {
    cnt[c + 1] += colptr[c + 1] - colptr[c];
    for (Ptr p = colptr[c] - baseval; p < colptr[c + 1] - baseval; ++p) {
      Idx r = rowind[p] - baseval;  // 0-based
      if (r > c) cnt[r + 1]++;
    }
  }
  g.colptr.assign(n + 1, 0);
  g.colptr[0] = 1;
  for (Int c = 0; c < n; ++c) g.colptr[c + 1] = g.colptr[c] + cnt[c + 1];
  g.rowind.assign(g.colptr[n] - 1, 0);
  std::vector<Ptr> pos(n);
  for (Int c = 0; c < n; ++c) pos[c] = g.colptr[c] - 1;
  // mirrored (upper) entries first, in increasing source-column order
  for (Int c = 0; c < n; ++c)
    for (Ptr p = colptr[c] - baseval; p < colptr[c + 1] - baseval; ++p) {
      Idx r = rowind[p] - baseval;
      if (r > c) g.rowind[pos[r]++] = c + 1;
    }
  for (Int c = 0; c < n; ++c)
    for (Ptr p = colptr[c] - baseval; p < colptr[c + 1] - baseval; ++p)
      g.rowind[pos[c]++] = rowind[p] - baseval + 1;
  // a column holds its mirrored rows (< its own index, ascending) and then
  // its own lower rows: sorted as soon as the input column is
  for (Int c = 0; c < n; ++c) {
    auto b = g.rowind.begin() + (g.colptr[c] - 1), e = g.rowind.begin() + (g.colptr[c + 1] - 1);
    if (!std::is_sorted(b, e)) std::sort(b, e);
  }
  g.keep_diag = true;
  return g;
}

// SparseMatrixGraph::SetKeepDiag (cpp:39-97).
void set_keep_diag(Graph& g, bool keep) {
  if (keep == g.keep_diag) return;
  Int n = g.n;
  if (keep) {
    std::vector<Idx> nr;
    nr.reserve(g.rowind.size() + n);
    std::vector<Ptr> nc(n + 1);
    for (Int lc = 0; lc < n; ++lc) {
      Idx col = lc + 1;
      Ptr b = g.colptr[lc] - 1, e = g.colptr[lc + 1] - 1;
      nc[lc] = (Ptr)nr.size() + 1;
      bool added = false;
      for (Ptr p = b; p < e; ++p) {
        Idx row = g.rowind[p];
        if (row < col) nr.push_back(row);
        if (row > col) {
          if (!added) {
            nr.push_back(col);
            added = true;
          }
          nr.push_back(row);
        }
      }
      if (!added) nr.push_back(col);
    }
    nc[n] = (Ptr)nr.size() + 1;
    g.colptr.swap(nc);
    g.rowind.swap(nr);
  } else {
    Ptr pos = 0;
    for (Int lc = 0; lc < n; ++lc) {
      Idx col = lc + 1;
      Ptr b = g.colptr[lc] - 1, e = g.colptr[lc + 1] - 1;
      g.colptr[lc] = pos + 1;
      for (Ptr p = b; p < e; ++p)
        if (g.rowind[p] != col) g.rowind[pos++] = g.rowind[p];
    }
    g.rowind.resize(pos);
    g.colptr[n] = pos + 1;
  }
  g.keep_diag = keep;
}

// DistSparseMatrixGraph::permute_ (cpp:508-640) for an expanded graph on one
// rank: entry (row,col) moves to (invp[row],invp[col]); columns re-sorted.
Graph permute_graph(const Graph& g, const std::vector<Int>& invp) {
  Graph p;
  p.n = g.n;
  p.keep_diag = g.keep_diag;
  Int n = g.n;
  std::vector<Ptr> cnt(n + 1, 0);
  for (Int c = 0; c < n; ++c) cnt[invp[c]] += g.colptr[c + 1] - g.colptr[c];
  p.colptr.assign(n + 1, 1);
  for (Int c = 0; c < n; ++c) p.colptr[c + 1] = p.colptr[c] + cnt[c + 1];
  p.rowind.assign(g.rowind.size(), 0);
  std::vector<Ptr> pos(n);
  for (Int c = 0; c < n; ++c) pos[c] = p.colptr[c] - 1;
  for (Int c = 0; c < n; ++c) {
    Int pc = invp[c] - 1;
    for (Ptr q = g.colptr[c] - 1; q < g.colptr[c + 1] - 1; ++q)
      p.rowind[pos[pc]++] = invp[g.rowind[q] - 1];
  }
  for (Int c = 0; c < n; ++c)
    std::sort(p.rowind.begin() + (p.colptr[c] - 1), p.rowind.begin() + (p.colptr[c + 1] - 1));
  return p;
}

// ---------------------------------------------------------------------------
// Elimination tree (src/ETree.cpp)
// ---------------------------------------------------------------------------
namespace {
// Liu's algorithm with path compression; the column loop is skipped when the
// column holds fewer than two entries (`jstrt < jstop`, cpp:271,345).
template <class RowOf>
void etree_liu(Int n, std::vector<Int>& parent, RowOf&& column) {
  parent.assign(n, 0);
  std::vector<Int> ancstr(n, 0);
  for (Int i = 1; i <= n; ++i) {
    parent[i - 1] = 0;
    ancstr[i - 1] = 0;
    const Idx* rows;
    Ptr cnt;
    column(i, rows, cnt);
    if (cnt < 2) continue;
    for (Ptr j = 0; j < cnt; ++j) {
      Int nbr = rows[j];
      if (nbr >= i) continue;
      if (ancstr[nbr - 1] == i) continue;
      bool brk = false;
      while (ancstr[nbr - 1] > 0) {
        if (ancstr[nbr - 1] == i) {
          brk = true;
          break;
        }
        Int next = ancstr[nbr - 1];
        ancstr[nbr - 1] = i;
        nbr = next;
      }
      if (!brk) {
        parent[nbr - 1] = i;
        ancstr[nbr - 1] = i;
      }
    }
  }
}
}  // namespace

void ETree::construct_serial(const Graph& sg, const Ordering& ord) {
  postordered = false;
  n = sg.n;
  std::vector<Idx> buf;
  etree_liu(n, parent, [&](Int i, const Idx*& rows, Ptr& cnt) {
    Int node = ord.perm[i - 1];
    Ptr b = sg.colptr[node - 1] - 1, e = sg.colptr[node] - 1;
    cnt = e - b;
    buf.resize(cnt);
    for (Ptr j = 0; j < cnt; ++j) buf[j] = ord.invp[sg.rowind[b + j] - 1];
    rows = buf.data();
  });
}

void ETree::construct_permuted(const Graph& pg) {
  postordered = false;
  n = pg.n;
  etree_liu(n, parent, [&](Int i, const Idx*& rows, Ptr& cnt) {
    Ptr b = pg.colptr[i - 1] - 1, e = pg.colptr[i] - 1;
    cnt = e - b;
    rows = pg.rowind.data() + b;
  });
}

namespace {
// ETree::BTreeToPO (cpp:25-56): depth-first traversal of the first-son /
// brother binary tree rooted at vertex n.
void btree_to_po(Int n, const std::vector<Int>& fson, const std::vector<Int>& brother,
                 std::vector<Int>& invpos) {
  std::vector<Int> stack(n);
  invpos.assign(n, 0);
  Int stacktop = 0, vertex = n, m = 0;
  bool exit = false;
  while (m < n) {
    do {
      stacktop++;
      stack[stacktop - 1] = vertex;
      vertex = fson[vertex - 1];
    } while (vertex > 0);
    while (vertex == 0) {
      if (stacktop <= 0) {
        exit = true;
        break;
      }
      vertex = stack[stacktop - 1];
      stacktop--;
      m++;
      invpos[vertex - 1] = m;
      vertex = brother[vertex - 1];
    }
    if (exit) break;
  }
}
}  // namespace

// ETree::PostOrderTree (cpp:59-119).
void ETree::postorder(Ordering& ord) {
  if (n <= 0 || postordered) return;
  std::vector<Int> fson(n, 0);
  std::vector<Int>& brother = poparent;
  brother.assign(n, 0);
  Int lroot = n;
  for (Int vertex = n - 1; vertex > 0; vertex--) {
    Int cur = parent[vertex - 1];
    if (cur == 0 || cur == vertex) {
      brother[lroot - 1] = vertex;
      lroot = vertex;
    } else {
      brother[vertex - 1] = fson[cur - 1];
      fson[cur - 1] = vertex;
    }
  }
  std::vector<Int> invpos;
  btree_to_po(n, fson, brother, invpos);
  for (Int i = 1; i <= n; i++) {
    Int nunode = invpos[i - 1];
    Int ndpar = parent[i - 1];
    if (ndpar > 0) ndpar = invpos[ndpar - 1];
    poparent[nunode - 1] = ndpar;
  }
  ord.compose(invpos);
  postordered = true;
}

// ETree::SortChildren (cpp:122-225): re-postorder so that the children of
// each node are visited in increasing column-count order, permuting cc.
void ETree::sort_children(std::vector<Int>& cc, Ordering& ord) {
  if (!postordered) postorder(ord);
  std::vector<Int> fson(n, 0), brother(n, 0), lson(n, 0);
  Int lroot = n;
  for (Int vertex = n - 1; vertex > 0; vertex--) {
    Int cur = PostParent(vertex - 1);
    if (cur == 0 || cur == vertex) {
      brother[lroot - 1] = vertex;
      lroot = vertex;
    } else {
      Int ndlson = lson[cur - 1];
      if (ndlson > 0) {
        if (cc[vertex - 1] >= cc[ndlson - 1]) {
          brother[vertex - 1] = fson[cur - 1];
          fson[cur - 1] = vertex;
        } else {
          brother[ndlson - 1] = vertex;
          lson[cur - 1] = vertex;
        }
      } else {
        fson[cur - 1] = vertex;
        lson[cur - 1] = vertex;
      }
    }
  }
  brother[lroot - 1] = 0;

  std::vector<Int> invpos;
  btree_to_po(n, fson, brother, invpos);
  for (Int i = 1; i <= n; i++) {
    Int nunode = invpos[i - 1];
    Int ndpar = poparent[i - 1];
    if (ndpar > 0) ndpar = invpos[ndpar - 1];
    brother[nunode - 1] = ndpar;
  }
  poparent = brother;
  std::vector<Int> tmp(n);
  for (Int node = 1; node <= n; ++node) tmp[invpos[node - 1] - 1] = cc[node - 1];
  cc = tmp;
  ord.compose(invpos);
}

}  // namespace spk
