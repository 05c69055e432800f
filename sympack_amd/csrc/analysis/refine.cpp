// Reordering of the columns inside each supernode by partition refinement,
// so that the row structures of the supernodes split into fewer runs of
// consecutive rows: symPACKMatrixMeta::refineSupernodes
// (include/sympack/impl/symPACKMatrixBase_impl.hpp:1006-1080), which calls
// the Ng-Peyton routine `ordsup` (src/ordsup.f) with an identity ordering and
// composes the result into the current one (Ordering::Compose,
// src/Ordering.cpp:1301-1314).  The supernode partition does not change; only
// the numbering inside every supernode does.
//
// The steps follow the reference routine one for one (which set is split,
// where the new half goes, the order of the free sets and of the processing),
// because the resulting permutation must be the same, index for index.  The
// data structures are this file's own: intrusive doubly linked lists over set
// and node ids and a binary min-heap with the same tie-breaking.
#include <utility>
#include <vector>

#include "analysis.hpp"

namespace spk {
namespace {

// Doubly linked list over the ids 1..n (0 = none); push_front and unlink.
struct Links {
  std::vector<int> next, prev;
  int head = 0;
  explicit Links(int n) : next(n + 1, 0), prev(n + 1, 0) {}
  void push_front(int x) {
    next[x] = head;
    if (head > 0) prev[head] = x;
    prev[x] = 0;
    head = x;
  }
  void unlink(int x) {
    const int a = next[x], b = prev[x];
    if (b == 0) head = a;
    else next[b] = a;
    if (a > 0) prev[a] = b;
  }
  int pop_front() {
    const int x = head;
    head = next[x];
    if (head > 0) prev[head] = 0;
    return x;
  }
};

// Min-heap of (key, vertex) with positions 1..size, as the reference's heap
// routines keep it (src/ins_heap.f, del_heap.f, mod_heap.f, move_up.f,
// move_down.f): a child rises while its key is strictly below its parent's, a
// parent sinks to the smaller child (the left one on a tie) while its key is
// strictly above either child's.
struct KeyHeap {
  std::vector<int> key, vtx, pos;  // key/vtx by heap position, pos by vertex
  int size = 0;
  explicit KeyHeap(int nv) : key(nv + 2, 0), vtx(nv + 2, 0), pos(nv + 1, 0) {}
  void swap_at(int a, int b) {
    std::swap(key[a], key[b]);
    std::swap(vtx[a], vtx[b]);
    pos[vtx[a]] = a;
    pos[vtx[b]] = b;
  }
  void rise(int i) {
    while (i > 1 && key[i] < key[i / 2]) {
      swap_at(i, i / 2);
      i /= 2;
    }
  }
  void sink(int i) {
    const int inf = 2000000000;
    while (i <= size) {
      const int l = 2 * i, r = l + 1;
      const int kl = l <= size ? key[l] : inf, kr = r <= size ? key[r] : inf;
      if (key[i] <= kl && key[i] <= kr) break;
      const int c = kl <= kr ? l : r;
      swap_at(i, c);
      i = c;
    }
  }
  void insert(int v, int k) {
    ++size;
    key[size] = k;
    vtx[size] = v;
    pos[v] = size;
    if (size > 1) rise(size);
  }
  // the vertex at position i takes key k and moves to its place
  void rekey(int i, int k) {
    key[i] = k;
    if (i > 1 && key[i / 2] > k) {
      rise(i);
      return;
    }
    const int l = 2 * i, r = l + 1;
    if ((l <= size && key[l] < k) || (r <= size && key[r] < k)) sink(i);
  }
  void remove(int v) {
    if (size == 0) return;
    const int i = pos[v];
    pos[v] = 0;
    if (i < size) {
      const int last = vtx[size], lk = key[size];
      vtx[i] = last;
      pos[last] = i;
      --size;
      rekey(i, lk);
    } else {
      --size;
    }
  }
};

// Processing order of the supernodes (src/mcssup.f): a supernode is taken
// after its parent; among the available ones, the heap gives the smallest
// key, -(off-diagonal row count) for ordflag 2, -(descendants in the
// supernodal etree, itself included) for ordflag 3.  order[k] (k = nsuper
// down to 1) is the k-th supernode taken.
void supernode_order(int ordflag, int nsuper, const std::vector<Int>& xsuper,
                     const std::vector<Int>& snode, const std::vector<Ptr>& xlindx,
                     const std::vector<Idx>& lindx, std::vector<int>& order) {
  std::vector<int> par(nsuper + 1, 0), ndesc(nsuper + 1, 1);
  for (int j = 1; j <= nsuper; ++j) {
    const int w = xsuper[j] - xsuper[j - 1];
    const Ptr first_off = xlindx[j - 1] + w;
    if (first_off <= xlindx[j] - 1) {
      par[j] = snode[lindx[first_off - 1] - 1];
      ndesc[par[j]] += ndesc[j];
    }
  }
  // children lists in increasing order (src/betree.f: built from the last
  // supernode down)
  std::vector<int> first_child(nsuper + 1, 0), sibling(nsuper + 1, 0);
  int last_root = nsuper;
  for (int j = nsuper - 1; j >= 1; --j) {
    if (par[j] <= 0 || par[j] == j) {
      sibling[last_root] = j;
      last_root = j;
    } else {
      sibling[j] = first_child[par[j]];
      first_child[par[j]] = j;
    }
  }
  if (nsuper >= 1) sibling[last_root] = 0;

  KeyHeap h(nsuper);
  for (int j = 1; j <= nsuper; ++j)
    if (par[j] == 0) h.insert(j, 0);
  order.assign(nsuper + 1, 0);
  for (int k = nsuper; k > 0; --k) {
    const int j = h.vtx[1];
    h.remove(j);
    for (int c = first_child[j]; c > 0; c = sibling[c]) {
      int key = 0;
      if (ordflag == 2) {
        const int w = xsuper[c] - xsuper[c - 1];
        key = -(int)(xlindx[c] - xlindx[c - 1] - w);
      } else {
        key = -ndesc[c];
      }
      h.insert(c, key);
    }
    order[k] = j;
  }
}

}  // namespace

// src/ordsup.f: sets of nodes (columns), initially one per supernode, kept in
// a list in which every supernode comes before its parent.  Supernodes are
// visited from the top of the processing order down; the nodes adjacent to
// the visited supernode (its off-diagonal rows) leave their set for a
// companion set, and each split set is replaced in the list by the two
// halves, adjacent half first or second (altflag 1 alternates this within a
// run of sets of one supernode, starting with "after").  Finally the nodes
// are renumbered inside their supernodes in list order.
// Returns invp2 (1-based): new position of each current node.
std::vector<Int> refine_supernodes(int ordflag, int altflag, Int n, const std::vector<Int>& xsuper,
                                   const std::vector<Int>& snode, const std::vector<Ptr>& xlindx,
                                   const std::vector<Idx>& lindx) {
  const int nsuper = (int)xsuper.size() - 1;
  std::vector<int> order;
  if (ordflag == 1) {
    order.assign(nsuper + 1, 0);
    for (int s = 1; s <= nsuper; ++s) order[s] = s;
  } else {
    supernode_order(ordflag, nsuper, xsuper, snode, xlindx, lindx, order);
  }

  std::vector<int> mark(n + 1, 0), set_of(n + 1, 0), comp(n + 1, 0), owner(n + 1, 0);
  for (int i = 1; i <= n; ++i) set_of[i] = snode[i - 1];
  Links free_sets(n), sets(n), runs(n);
  std::vector<int> node_head(n + 1, 0);
  Links nodes(n);  // node lists, one per set: heads live in node_head
  auto node_push = [&](int s, int v) {
    nodes.head = node_head[s];
    nodes.push_front(v);
    node_head[s] = nodes.head;
  };
  auto node_unlink = [&](int s, int v) {
    nodes.head = node_head[s];
    nodes.unlink(v);
    node_head[s] = nodes.head;
  };
  auto singleton = [&](int s) { return nodes.next[node_head[s]] == 0; };

  for (int s = nsuper + 1; s <= n; ++s) free_sets.push_front(s);
  for (int k = nsuper; k >= 1; --k) {
    sets.push_front(order[k]);
    owner[order[k]] = order[k];
  }
  for (int k = 1; k <= nsuper; ++k) {
    const int s = order[k];
    for (int v = xsuper[s] - 1; v >= xsuper[s - 1]; --v) node_push(s, v);
    if (singleton(s)) comp[s] = -1;
  }

  for (int k = nsuper; k >= 1; --k) {
    const int cur = order[k];
    runs.head = 0;
    // ---- move the adjacent nodes to the companion sets; record the runs of
    // consecutive marked sets of one supernode by their first set ----
    const Ptr beg = xlindx[cur - 1] + (xsuper[cur] - xsuper[cur - 1]), end = xlindx[cur] - 1;
    for (Ptr p = beg; p <= end; ++p) {
      const int v = lindx[p - 1];
      const int s = set_of[v];
      if (comp[s] == 0) comp[s] = free_sets.pop_front();
      if (mark[s] == 0) {
        const int after = sets.next[s], before = sets.prev[s];
        if (after != 0 && mark[after] == 1 && owner[after] == owner[s]) runs.unlink(after);
        if (mark[before] == 0 || owner[before] != owner[s]) runs.push_front(s);
        mark[s] = 1;
      }
      if (comp[s] > 0) {
        node_unlink(s, v);
        node_push(comp[s], v);
        set_of[v] = comp[s];
      }
    }
    // ---- split the marked sets, run by run ----
    int up = 0;
    for (int s = runs.head; s > 0; s = runs.next[s]) {
      if (altflag == 0) up = 0;
      int prev_owner = 0;
      int t = s;
      while (t > 0 && mark[t] == 1) {
        const int o = owner[t];
        if (altflag == 1 && o != prev_owner) up = 1;
        prev_owner = o;
        const int before = sets.prev[t], after = sets.next[t];
        if (comp[t] != -1) {
          const int tp = comp[t];
          comp[t] = 0;
          if (node_head[t] == 0) {
            // every node of t moved: tp takes t's place, t is free again
            free_sets.push_front(t);
            if (before == 0) sets.head = tp;
            else sets.next[before] = tp;
            sets.prev[tp] = before;
            sets.next[tp] = after;
            if (after > 0) sets.prev[after] = tp;
            owner[tp] = owner[t];
            owner[t] = 0;
          } else {
            if (up == 1) {  // t, tp
              sets.next[t] = tp;
              sets.prev[tp] = t;
              sets.next[tp] = after;
              if (after > 0) sets.prev[after] = tp;
              if (altflag == 1) up = 0;
            } else {  // tp, t
              if (before == 0) sets.head = tp;
              else sets.next[before] = tp;
              sets.prev[tp] = before;
              sets.next[tp] = t;
              sets.prev[t] = tp;
              if (altflag == 1) up = 1;
            }
            if (singleton(t)) comp[t] = -1;
            if (singleton(tp)) comp[tp] = -1;
            owner[tp] = owner[t];
          }
        }
        mark[t] = 0;
        t = after;
      }
    }
  }

  // ---- number the nodes of every supernode in list order ----
  std::vector<Int> invp2(n, 0);
  std::vector<Int> nextnum(nsuper + 1, 0);
  for (int s = 1; s <= nsuper; ++s) nextnum[s] = xsuper[s - 1];
  for (int s = sets.head; s > 0; s = sets.next[s])
    for (int v = node_head[s]; v > 0; v = nodes.next[v]) invp2[v - 1] = nextnum[snode[v - 1]]++;
  return invp2;
}

}  // namespace spk
