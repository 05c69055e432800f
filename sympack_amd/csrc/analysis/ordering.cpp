// Fill-reducing orderings used by the supernodal path.
//
// These are restatements of the SPARSPAK-style Fortran routines that symPACK
// links (src/ordmmd.F, genmmd.F, mmdint.F, mmdelm.F, mmdupd.F, mmdnum.F,
// genrcm.F, rcm.F, fnroot.F, rootls.F, degree.F, boxnd.f, gridnd.f) and of
// the C++ front-ends in src/Ordering.cpp.  The permutations must be
// identical to the reference's, so the control flow (including the order in
// which degree lists are scanned and ties broken) follows the Fortran
// statement by statement.  All work arrays are 1-based: element 0 is unused.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "analysis.hpp"

namespace spk {

// ---------------------------------------------------------------------------
// Multiple minimum degree
// ---------------------------------------------------------------------------
namespace {

// mmdint.F: degree lists keyed by (degree+1), qsize=1, markers cleared.
void mmd_init(Int neqns, const Int* xadj, Int* dhead, Int* dforw, Int* dbakw, Int* qsize,
              Int* llist, Int* marker) {
  for (Int node = 1; node <= neqns; ++node) {
    dhead[node] = 0;
    qsize[node] = 1;
    marker[node] = 0;
    llist[node] = 0;
  }
  for (Int node = 1; node <= neqns; ++node) {
    Int ndeg = xadj[node + 1] - xadj[node] + 1;
    Int fnode = dhead[ndeg];
    dforw[node] = fnode;
    dhead[ndeg] = node;
    if (fnode > 0) dbakw[fnode] = node;
    dbakw[node] = -ndeg;
  }
}

// mmdelm.F: eliminate mdnode, build its reachable set in place and update
// the quotient graph (merging nodes that lose every active neighbour).
void mmd_eliminate(Int mdnode, const Int* xadj, Int* adjncy, Int* dhead, Int* dforw,
                   Int* dbakw, Int* qsize, Int* llist, Int* marker, Int maxint, Int tag) {
  Int elmnt, i, istop, istrt, j, jstop, jstrt, link, nabor, node, npv, nqnbrs, nxnode,
      pvnode, rlmt, rloc, rnode, xqnbr, slink;

  marker[mdnode] = tag;
  istrt = xadj[mdnode];
  istop = xadj[mdnode + 1] - 1;
  elmnt = 0;
  rloc = istrt;
  rlmt = istop;
  for (i = istrt; i <= istop; ++i) {
    nabor = adjncy[i];
    if (nabor == 0) break;
    if (marker[nabor] >= tag) continue;
    marker[nabor] = tag;
    if (dforw[nabor] < 0) {
      llist[nabor] = elmnt;
      elmnt = nabor;
    } else {
      adjncy[rloc] = nabor;
      rloc = rloc + 1;
    }
  }
  // merge with reachable nodes from generalized elements.
  while (elmnt > 0) {
    adjncy[rlmt] = -elmnt;
    link = elmnt;
  walk_element:
    jstrt = xadj[link];
    jstop = xadj[link + 1] - 1;
    for (j = jstrt; j <= jstop; ++j) {
      node = adjncy[j];
      link = -node;
      if (node < 0) goto walk_element;
      if (node == 0) break;
      if (marker[node] >= tag || dforw[node] < 0) continue;
      marker[node] = tag;
      // use storage from eliminated nodes if necessary.
      while (rloc >= rlmt) {
        slink = -adjncy[rlmt];
        rloc = xadj[slink];
        rlmt = xadj[slink + 1] - 1;
      }
      adjncy[rloc] = node;
      rloc = rloc + 1;
    }
    elmnt = llist[elmnt];
  }
  if (rloc <= rlmt) adjncy[rloc] = 0;

  // for each node in the reachable set ...
  link = mdnode;
walk_reach:
  istrt = xadj[link];
  istop = xadj[link + 1] - 1;
  for (i = istrt; i <= istop; ++i) {
    rnode = adjncy[i];
    link = -rnode;
    if (rnode < 0) goto walk_reach;
    if (rnode == 0) return;
    // if rnode is in the degree list structure, remove it.
    pvnode = dbakw[rnode];
    if (!(pvnode == 0 || pvnode == -maxint)) {
      nxnode = dforw[rnode];
      if (nxnode > 0) dbakw[nxnode] = pvnode;
      if (pvnode > 0) dforw[pvnode] = nxnode;
      npv = -pvnode;
      if (pvnode < 0) dhead[npv] = nxnode;
    }
    // purge inactive quotient nabors of rnode.
    jstrt = xadj[rnode];
    jstop = xadj[rnode + 1] - 1;
    xqnbr = jstrt;
    for (j = jstrt; j <= jstop; ++j) {
      nabor = adjncy[j];
      if (nabor == 0) break;
      if (marker[nabor] >= tag) continue;
      adjncy[xqnbr] = nabor;
      xqnbr = xqnbr + 1;
    }
    nqnbrs = xqnbr - jstrt;
    if (nqnbrs <= 0) {
      // no active nabor: merge rnode with mdnode.
      qsize[mdnode] = qsize[mdnode] + qsize[rnode];
      qsize[rnode] = 0;
      marker[rnode] = maxint;
      dforw[rnode] = -mdnode;
      dbakw[rnode] = -maxint;
      continue;
    }
    // flag rnode for degree update, add mdnode as a nabor of rnode.
    dforw[rnode] = nqnbrs + 1;
    dbakw[rnode] = 0;
    adjncy[xqnbr] = mdnode;
    xqnbr = xqnbr + 1;
    if (xqnbr <= jstop) adjncy[xqnbr] = 0;
  }
}

// mmdupd.F: recompute external degrees of the nodes touched by the
// elements eliminated in this pass (two-nabor fast path and general path).
void mmd_update(Int ehead, Int neqns, const Int* xadj, Int* adjncy, Int delta, Int& mdeg,
                Int* dhead, Int* dforw, Int* dbakw, Int* qsize, Int* llist, Int* marker,
                Int maxint, Int& tag) {
  Int deg, deg0, elmnt, enode, fnode, i, iq2, istop, istrt, j, jstop, jstrt, link, mdeg0,
      mtag, nabor, node, q2head, qxhead;

  mdeg0 = mdeg + delta;
  elmnt = ehead;
next_element:
  if (elmnt <= 0) return;
  mtag = tag + mdeg0;
  if (mtag >= maxint) {
    tag = 1;
    for (i = 1; i <= neqns; ++i)
      if (marker[i] < maxint) marker[i] = 0;
    mtag = tag + mdeg0;
  }
  // two lists: nodes with two nabors (q2head) and more (qxhead).
  q2head = 0;
  qxhead = 0;
  deg0 = 0;
  link = elmnt;
scan_element:
  istrt = xadj[link];
  istop = xadj[link + 1] - 1;
  for (i = istrt; i <= istop; ++i) {
    enode = adjncy[i];
    link = -enode;
    if (enode < 0) goto scan_element;
    if (enode == 0) break;
    if (qsize[enode] == 0) continue;
    deg0 = deg0 + qsize[enode];
    marker[enode] = mtag;
    if (dbakw[enode] != 0) continue;
    if (dforw[enode] == 2) {
      llist[enode] = q2head;
      q2head = enode;
    } else {
      llist[enode] = qxhead;
      qxhead = enode;
    }
  }

  // q2 list.
  enode = q2head;
  iq2 = 1;
q2_loop:
  if (enode <= 0) goto qx_start;
  if (dbakw[enode] != 0) goto next_enode;
  tag = tag + 1;
  deg = deg0;
  istrt = xadj[enode];
  nabor = adjncy[istrt];
  if (nabor == elmnt) nabor = adjncy[istrt + 1];
  link = nabor;
  if (dforw[nabor] >= 0) {
    deg = deg + qsize[nabor];
    goto set_degree;
  }
q2_walk:
  istrt = xadj[link];
  istop = xadj[link + 1] - 1;
  for (i = istrt; i <= istop; ++i) {
    node = adjncy[i];
    link = -node;
    if (node == enode) continue;
    if (node < 0) goto q2_walk;
    if (node == 0) goto set_degree;
    if (qsize[node] == 0) continue;
    if (marker[node] < tag) {
      marker[node] = tag;
      deg = deg + qsize[node];
      continue;
    }
    // node indistinguishable from enode: merge.
    if (dbakw[node] != 0) continue;
    if (dforw[node] == 2) {
      qsize[enode] = qsize[enode] + qsize[node];
      qsize[node] = 0;
      marker[node] = maxint;
      dforw[node] = -enode;
      dbakw[node] = -maxint;
      continue;
    }
    // node outmatched by enode.
    if (dbakw[node] == 0) dbakw[node] = -maxint;
  }
  goto set_degree;

qx_start:
  enode = qxhead;
  iq2 = 0;
qx_loop:
  if (enode <= 0) goto element_done;
  if (dbakw[enode] != 0) goto next_enode;
  tag = tag + 1;
  deg = deg0;
  istrt = xadj[enode];
  istop = xadj[enode + 1] - 1;
  for (i = istrt; i <= istop; ++i) {
    nabor = adjncy[i];
    if (nabor == 0) break;
    if (marker[nabor] >= tag) continue;
    marker[nabor] = tag;
    link = nabor;
    if (dforw[nabor] >= 0) {
      deg = deg + qsize[nabor];
      continue;
    }
  qx_walk:
    jstrt = xadj[link];
    jstop = xadj[link + 1] - 1;
    for (j = jstrt; j <= jstop; ++j) {
      node = adjncy[j];
      link = -node;
      if (node < 0) goto qx_walk;
      if (node == 0) break;
      if (marker[node] >= tag) continue;
      marker[node] = tag;
      deg = deg + qsize[node];
    }
  }

set_degree:
  deg = deg - qsize[enode] + 1;
  fnode = dhead[deg];
  dforw[enode] = fnode;
  dbakw[enode] = -deg;
  if (fnode > 0) dbakw[fnode] = enode;
  dhead[deg] = enode;
  if (deg < mdeg) mdeg = deg;
next_enode:
  enode = llist[enode];
  if (iq2 == 1) goto q2_loop;
  goto qx_loop;

element_done:
  tag = mtag;
  elmnt = llist[elmnt];
  goto next_element;
}

// mmdnum.F: number merged nodes after their representatives.
void mmd_number(Int neqns, Int* perm, Int* invp, const Int* qsize) {
  Int father, nextf, node, nqsize, num, root;
  for (node = 1; node <= neqns; ++node) {
    nqsize = qsize[node];
    if (nqsize <= 0) perm[node] = invp[node];
    if (nqsize > 0) perm[node] = -invp[node];
  }
  for (node = 1; node <= neqns; ++node) {
    if (perm[node] > 0) continue;
    father = node;
    while (perm[father] <= 0) father = -perm[father];
    root = father;
    num = perm[root] + 1;
    invp[node] = -num;
    perm[root] = num;
    father = node;
    for (;;) {
      nextf = -perm[father];
      if (nextf <= 0) break;
      perm[father] = -root;
      father = nextf;
    }
  }
  for (node = 1; node <= neqns; ++node) {
    num = -invp[node];
    invp[node] = num;
    perm[num] = node;
  }
}

// genmmd.F with delta=0, maxint=2e9 (ordmmd.F:71-73).
void genmmd(Int neqns, Int* xadj, Int* adjncy, Int* invp, Int* perm, Int delta, Int* dhead,
            Int* qsize, Int* llist, Int* marker, Int maxint, Int& nofsub) {
  Int ehead, i, mdeg, mdlmt, mdnode, nextmd, num, tag;
  if (neqns <= 0) return;
  nofsub = 0;
  mmd_init(neqns, xadj, dhead, invp, perm, qsize, llist, marker);
  num = 1;
  // eliminate all isolated nodes.
  nextmd = dhead[1];
  while (nextmd > 0) {
    mdnode = nextmd;
    nextmd = invp[mdnode];
    marker[mdnode] = maxint;
    invp[mdnode] = -num;
    num = num + 1;
  }
  if (num > neqns) goto done;
  tag = 1;
  dhead[1] = 0;
  mdeg = 2;
search_min:
  while (dhead[mdeg] <= 0) mdeg = mdeg + 1;
  mdlmt = mdeg + delta;
  ehead = 0;
pick:
  mdnode = dhead[mdeg];
  if (mdnode <= 0) {
    mdeg = mdeg + 1;
    if (mdeg > mdlmt) goto update;
    goto pick;
  }
  // remove mdnode from the degree structure.
  nextmd = invp[mdnode];
  dhead[mdeg] = nextmd;
  if (nextmd > 0) perm[nextmd] = -mdeg;
  invp[mdnode] = -num;
  nofsub = nofsub + mdeg + qsize[mdnode] - 2;
  if (num + qsize[mdnode] > neqns) goto done;
  tag = tag + 1;
  if (tag >= maxint) {
    tag = 1;
    for (i = 1; i <= neqns; ++i)
      if (marker[i] < maxint) marker[i] = 0;
  }
  mmd_eliminate(mdnode, xadj, adjncy, dhead, invp, perm, qsize, llist, marker, maxint, tag);
  num = num + qsize[mdnode];
  llist[mdnode] = ehead;
  ehead = mdnode;
  if (delta >= 0) goto pick;
update:
  if (num > neqns) goto done;
  mmd_update(ehead, neqns, xadj, adjncy, delta, mdeg, dhead, invp, perm, qsize, llist, marker,
             maxint, tag);
  goto search_min;
done:
  mmd_number(neqns, perm, invp, qsize);
}

void check_graph_for_ordering(const Graph& g, const char* who) {
  if (g.keep_diag)
    throw std::logic_error(std::string("SparseMatrixGraph must be expanded and not including the "
                                       "diagonal in order to call ") +
                           who + "\n");
}

void perm_from_invp(Ordering& ord) {
  Int n = (Int)ord.invp.size();
  ord.perm.assign(n, 0);
  for (Int i = 1; i <= n; ++i) ord.perm[ord.invp[i - 1] - 1] = i;
}

}  // namespace

void order_mmd(const Graph& g, Ordering& ord) {
  check_graph_for_ordering(g, "MMD");
  Int n = g.n;
  // xadj/adjncy are destroyed by genmmd: work on 1-based copies.
  std::vector<Int> xadj(n + 2, 0), adjncy(g.rowind.size() + 2, 0);
  for (Int i = 0; i <= n; ++i) xadj[i + 1] = (Int)g.colptr[i];
  for (size_t i = 0; i < g.rowind.size(); ++i) adjncy[i + 1] = g.rowind[i];
  std::vector<Int> invp(n + 1, 0), perm(n + 1, 0);
  std::vector<Int> iwork(4 * (size_t)n + 4, 0);
  Int* dhead = iwork.data();
  Int* qsize = iwork.data() + n;
  Int* llist = iwork.data() + 2 * n;
  Int* marker = iwork.data() + 3 * n;
  Int nofsub = 0;
  genmmd(n, xadj.data(), adjncy.data(), invp.data(), perm.data(), 0, dhead, qsize, llist, marker,
         2000000000, nofsub);
  ord.invp.assign(invp.begin() + 1, invp.end());
  perm_from_invp(ord);  // Ordering.cpp:211-214
}

// ---------------------------------------------------------------------------
// Reverse Cuthill-McKee
// ---------------------------------------------------------------------------
namespace {

// rootls.F: rooted level structure of the masked component of root.
void rcm_rootls(Int root, const Int* xadj, const Int* adjncy, Int* mask, Int& nlvl, Int* xls,
                Int* ls) {
  Int i, j, jstop, jstrt, lbegin, ccsize, lvlend, lvsize, nbr, node;
  mask[root] = 0;
  ls[1] = root;
  nlvl = 0;
  lvlend = 0;
  ccsize = 1;
  do {
    lbegin = lvlend + 1;
    lvlend = ccsize;
    nlvl = nlvl + 1;
    xls[nlvl] = lbegin;
    for (i = lbegin; i <= lvlend; ++i) {
      node = ls[i];
      jstrt = xadj[node];
      jstop = xadj[node + 1] - 1;
      if (jstop < jstrt) continue;
      for (j = jstrt; j <= jstop; ++j) {
        nbr = adjncy[j];
        if (mask[nbr] != 0) {
          ccsize = ccsize + 1;
          ls[ccsize] = nbr;
          mask[nbr] = 0;
        }
      }
    }
    lvsize = ccsize - lvlend;
  } while (lvsize > 0);
  xls[nlvl + 1] = lvlend + 1;
  for (i = 1; i <= ccsize; ++i) mask[ls[i]] = 1;
}

// fnroot.F: pseudo-peripheral node search.
void rcm_fnroot(Int& root, const Int* xadj, const Int* adjncy, Int* mask, Int& nlvl, Int* xls,
                Int* ls) {
  Int ccsize, j, jstrt, k, kstop, kstrt, mindeg, nabor, ndeg, node, nunlvl;
  rcm_rootls(root, xadj, adjncy, mask, nlvl, xls, ls);
  ccsize = xls[nlvl + 1] - 1;
  if (nlvl == 1 || nlvl == ccsize) return;
  for (;;) {
    jstrt = xls[nlvl];
    mindeg = ccsize;
    root = ls[jstrt];
    if (ccsize != jstrt) {
      for (j = jstrt; j <= ccsize; ++j) {
        node = ls[j];
        ndeg = 0;
        kstrt = xadj[node];
        kstop = xadj[node + 1] - 1;
        for (k = kstrt; k <= kstop; ++k) {
          nabor = adjncy[k];
          if (mask[nabor] > 0) ndeg = ndeg + 1;
        }
        if (ndeg >= mindeg) continue;
        root = node;
        mindeg = ndeg;
      }
    }
    rcm_rootls(root, xadj, adjncy, mask, nunlvl, xls, ls);
    if (nunlvl <= nlvl) return;
    nlvl = nunlvl;
    if (nlvl >= ccsize) return;
  }
}

// degree.F: masked degrees of the component of root (xadj sign trick).
void rcm_degree(Int root, Int* xadj, const Int* adjncy, const Int* mask, Int* deg, Int& ccsize,
                Int* ls) {
  Int i, ideg, j, jstop, jstrt, lbegin, lvlend, lvsize, nbr, node;
  ls[1] = root;
  xadj[root] = -xadj[root];
  lvlend = 0;
  ccsize = 1;
  do {
    lbegin = lvlend + 1;
    lvlend = ccsize;
    for (i = lbegin; i <= lvlend; ++i) {
      node = ls[i];
      jstrt = -xadj[node];
      jstop = std::abs(xadj[node + 1]) - 1;
      ideg = 0;
      if (jstop >= jstrt) {
        for (j = jstrt; j <= jstop; ++j) {
          nbr = adjncy[j];
          if (mask[nbr] == 0) continue;
          ideg = ideg + 1;
          if (xadj[nbr] < 0) continue;
          xadj[nbr] = -xadj[nbr];
          ccsize = ccsize + 1;
          ls[ccsize] = nbr;
        }
      }
      deg[node] = ideg;
    }
    lvsize = ccsize - lvlend;
  } while (lvsize > 0);
  for (i = 1; i <= ccsize; ++i) {
    node = ls[i];
    xadj[node] = -xadj[node];
  }
}

// rcm.F: Cuthill-McKee from root, then reversed.
void rcm_component(Int root, Int* xadj, const Int* adjncy, Int* mask, Int* perm, Int& ccsize,
                   Int* deg) {
  Int fnbr, i, j, jstop, jstrt, k, l, lbegin, lnbr, lperm, lvlend, nbr, node;
  rcm_degree(root, xadj, adjncy, mask, deg, ccsize, perm);
  mask[root] = 0;
  if (ccsize <= 1) return;
  lvlend = 0;
  lnbr = 1;
  do {
    lbegin = lvlend + 1;
    lvlend = lnbr;
    for (i = lbegin; i <= lvlend; ++i) {
      node = perm[i];
      jstrt = xadj[node];
      jstop = xadj[node + 1] - 1;
      fnbr = lnbr + 1;
      for (j = jstrt; j <= jstop; ++j) {
        nbr = adjncy[j];
        if (mask[nbr] == 0) continue;
        lnbr = lnbr + 1;
        mask[nbr] = 0;
        perm[lnbr] = nbr;
      }
      if (fnbr >= lnbr) continue;
      // linear insertion sort of the new nodes by increasing degree.
      k = fnbr;
      do {
        l = k;
        k = k + 1;
        nbr = perm[k];
        while (l >= fnbr) {
          lperm = perm[l];
          if (deg[lperm] <= deg[nbr]) break;
          perm[l + 1] = lperm;
          l = l - 1;
        }
        perm[l + 1] = nbr;
      } while (k < lnbr);
    }
  } while (lnbr > lvlend);
  k = ccsize / 2;
  l = ccsize;
  for (i = 1; i <= k; ++i) {
    lperm = perm[l];
    perm[l] = perm[i];
    perm[i] = lperm;
    l = l - 1;
  }
}

}  // namespace

void order_rcm(const Graph& g, Ordering& ord) {
  check_graph_for_ordering(g, "RCM");
  Int n = g.n;
  std::vector<Int> xadj(n + 2, 0), adjncy(g.rowind.size() + 2, 0);
  for (Int i = 0; i <= n; ++i) xadj[i + 1] = (Int)g.colptr[i];
  for (size_t i = 0; i < g.rowind.size(); ++i) adjncy[i + 1] = g.rowind[i];
  std::vector<Int> perm(2 * (size_t)n + 2, 0), mask(2 * (size_t)n + 2, 0),
      xls(2 * (size_t)n + 2, 0);
  if (n > 0) {
    // genrcm.F
    for (Int i = 1; i <= n; ++i) mask[i] = 1;
    Int num = 1;
    for (Int i = 1; i <= n; ++i) {
      if (mask[i] == 0) continue;
      Int root = i, nlvl = 0, ccsize = 0;
      // PERM(NUM) is passed as the 1-based array start: shift by num-1.
      Int* pslice = perm.data() + (num - 1);
      rcm_fnroot(root, xadj.data(), adjncy.data(), mask.data(), nlvl, xls.data(), pslice);
      rcm_component(root, xadj.data(), adjncy.data(), mask.data(), pslice, ccsize, xls.data());
      num = num + ccsize;
      if (num > n) break;
    }
  }
  // Ordering.cpp:290-294: invp from perm.
  ord.perm.assign(perm.begin() + 1, perm.begin() + 1 + n);
  ord.invp.assign(n, 0);
  for (Int i = 1; i <= n; ++i) ord.invp[ord.perm[i - 1] - 1] = i;
  // Ordering.cpp:321-325 then rebuilds perm from invp (same thing).
  perm_from_invp(ord);
}

// ---------------------------------------------------------------------------
// Geometric nested dissection (boxnd.f / gridnd.f)
// ---------------------------------------------------------------------------
int boxnd(Int p, Int q, Int r, std::vector<Int>& box) {
  // BOX(P,Q,*) is column-major: node (ii,jj,kk) -> (ii-1) + P*(jj-1) + P*Q*(kk-1).
  box.assign((size_t)p * q * r, 0);
  auto at = [&](Int ii, Int jj, Int kk) -> Int& {
    return box[(size_t)(ii - 1) + (size_t)p * (jj - 1) + (size_t)p * q * (kk - 1)];
  };
  Int last = p * q * r;
  std::vector<Int> stack;
  stack.reserve(64 * 6);
  stack.insert(stack.end(), {1, 1, 1, p, q, r});
  while (!stack.empty()) {
    size_t top = stack.size();
    Int i = stack[top - 6], j = stack[top - 5], k = stack[top - 4];
    Int m = stack[top - 3], n = stack[top - 2], l = stack[top - 1];
    stack.resize(top - 6);
    auto label_all = [&]() {
      for (Int ii = i + m - 1; ii >= i; --ii)
        for (Int jj = j + n - 1; jj >= j; --jj)
          for (Int kk = k + l - 1; kk >= k; --kk) at(ii, jj, kk) = last--;
    };
    if (m >= n && m >= l) {
      if (m <= 2) {
        label_all();
      } else {
        Int half = (m - 1) / 2, mid = i + half;
        stack.insert(stack.end(), {i, j, k, half, n, l});
        stack.insert(stack.end(), {mid + 1, j, k, m - half - 1, n, l});
        for (Int jj = j + n - 1; jj >= j; --jj)
          for (Int kk = k + l - 1; kk >= k; --kk) at(mid, jj, kk) = last--;
      }
    } else if (n >= m && n >= l) {
      if (n <= 2) {
        label_all();
      } else {
        Int half = (n - 1) / 2, mid = j + half;
        stack.insert(stack.end(), {i, j, k, m, half, l});
        stack.insert(stack.end(), {i, mid + 1, k, m, n - half - 1, l});
        for (Int ii = i + m - 1; ii >= i; --ii)
          for (Int kk = k + l - 1; kk >= k; --kk) at(ii, mid, kk) = last--;
      }
    } else {
      if (l <= 2) {
        label_all();
      } else {
        Int half = (l - 1) / 2, mid = k + half;
        stack.insert(stack.end(), {i, j, k, m, n, half});
        stack.insert(stack.end(), {i, j, mid + 1, m, n, l - half - 1});
        for (Int ii = i + m - 1; ii >= i; --ii)
          for (Int jj = j + n - 1; jj >= j; --jj) at(ii, jj, mid) = last--;
      }
    }
  }
  return 0;
}

int gridnd(Int p, Int q, std::vector<Int>& grid) {
  grid.assign((size_t)p * q, 0);
  auto at = [&](Int ii, Int jj) -> Int& { return grid[(size_t)(ii - 1) + (size_t)p * (jj - 1)]; };
  Int last = p * q;
  std::vector<Int> stack;
  stack.insert(stack.end(), {1, 1, p, q});
  while (!stack.empty()) {
    size_t top = stack.size();
    Int i = stack[top - 4], j = stack[top - 3], m = stack[top - 2], n = stack[top - 1];
    stack.resize(top - 4);
    auto label_all = [&]() {
      for (Int ii = i + m - 1; ii >= i; --ii)
        for (Int jj = j + n - 1; jj >= j; --jj) at(ii, jj) = last--;
    };
    if (m >= n) {
      if (m <= 2) {
        label_all();
      } else {
        Int half = (m - 1) / 2, mid = i + half;
        stack.insert(stack.end(), {i, j, half, n});
        stack.insert(stack.end(), {mid + 1, j, m - half - 1, n});
        for (Int jj = j + n - 1; jj >= j; --jj) at(mid, jj) = last--;
      }
    } else {
      if (n <= 2) {
        label_all();
      } else {
        Int half = (n - 1) / 2, mid = j + half;
        stack.insert(stack.end(), {i, j, m, half});
        stack.insert(stack.end(), {i, mid + 1, m, n - half - 1});
        for (Int ii = i + m - 1; ii >= i; --ii) at(ii, mid) = last--;
      }
    }
  }
  return 0;
}

void order_ndbox(Int n, Ordering& ord) {
  // Ordering::NDBOX (Ordering.cpp:344-374): k = ceil(n^(1/3)) and the box is k^3.
  Int k = (Int)std::ceil(std::pow((double)n, 1.0 / 3.0));
  if ((int64_t)k * k * k != n)
    throw std::logic_error("NDBOX ordering needs n to be a perfect cube (k^3 grid, k=" +
                           std::to_string(k) + ")");
  std::vector<Int> box;
  boxnd(k, k, k, box);
  ord.invp = box;
  perm_from_invp(ord);
}

void order_ndgrid(Int n, Ordering& ord) {
  // Ordering::NDGRID (Ordering.cpp:376-408): k = ceil(n^(1/2)) and the grid is k^2.
  Int k = (Int)std::ceil(std::pow((double)n, 1.0 / 2.0));
  if ((int64_t)k * k != n)
    throw std::logic_error("NDGRID ordering needs n to be a perfect square (k^2 grid, k=" +
                           std::to_string(k) + ")");
  std::vector<Int> grid;
  gridnd(k, k, grid);
  ord.invp = grid;
  perm_from_invp(ord);
}

// ---------------------------------------------------------------------------
// Ordering helpers (src/Ordering.cpp:1289-1315)
// ---------------------------------------------------------------------------
void Ordering::compose(const std::vector<Int>& invp2) {
  Int n = (Int)invp.size();
  for (Int i = 1; i <= n; ++i) {
    Int interm = invp[i - 1];
    invp[i - 1] = invp2[interm - 1];
  }
  for (Int i = 1; i <= n; ++i) perm[invp[i - 1] - 1] = i;
}

void Ordering::relative_invp(const std::vector<Int>& frominvp, std::vector<Int>& relinvp) const {
  Int n = (Int)invp.size();
  relinvp.assign(n, 0);
  for (Int node = 1; node <= n; ++node) {
    Int i = perm[node - 1];
    Int interm = frominvp[i - 1];
    relinvp[interm - 1] = node;
  }
}

}  // namespace spk
