// Column counts, supernode partition, relaxation and the relaxed symbolic
// factorization, restating include/sympack/impl/symPACKMatrixBase_impl.hpp
// and the driver sequence of symPACKMatrix2D::SymbolicFactorization
// (include/sympack/symPACKMatrix2D.hpp:2699-3081).
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

#include "analysis.hpp"

namespace spk {

// getLColRowCount(SparseMatrixGraph&) (impl:309-466): Gilbert-Ng-Peyton
// column/row counts on the postordered etree; sg holds the diagonal and is
// indexed in the original numbering (rows mapped through invp).
void lcol_row_count(const Graph& sg, const ETree& et, const Ordering& ord, std::vector<Int>& cc,
                    std::vector<Int>& rc) {
  Int size = sg.n;
  cc.assign(size, 0);
  rc.assign(size, 0);
  std::vector<Idx> level(size + 1), fdesc(size + 1), nchild(size + 1);
  std::vector<Int> weight(size + 1, 0);
  std::vector<Idx> set(size), prvlf(size), prvnbr(size);
  Idx xsup = 1;
  level[0] = 0;
  for (Idx k = size; k >= 1; --k) {
    rc[k - 1] = 1;
    cc[k - 1] = 0;
    set[k - 1] = k;
    prvlf[k - 1] = 0;
    prvnbr[k - 1] = 0;
    level[k] = level[et.PostParent(k - 1)] + 1;
    weight[k] = 1;
    fdesc[k] = k;
    nchild[k] = 0;
  }
  nchild[0] = 0;
  fdesc[0] = 0;
  for (Idx k = 1; k < size; ++k) {
    Idx parent = et.PostParent(k - 1);
    weight[parent] = 0;
    ++nchild[parent];
    Idx ifdesc = fdesc[k];
    if (ifdesc < fdesc[parent]) fdesc[parent] = ifdesc;
  }
  for (Idx lownbr = 1; lownbr <= size; ++lownbr) {
    Int lflag = 0;
    Idx ifdesc = fdesc[lownbr];
    Idx oldnbr = ord.perm[lownbr - 1];
    Ptr jstrt = sg.colptr[oldnbr - 1], jstop = sg.colptr[oldnbr] - 1;
    for (Ptr j = jstrt; j <= jstop; ++j) {
      Idx hinbr = ord.invp[sg.rowind[j - 1] - 1];
      if (hinbr > lownbr) {
        if (ifdesc > prvnbr[hinbr - 1]) {
          ++weight[lownbr];
          Idx pleaf = prvlf[hinbr - 1];
          if (pleaf == 0) {
            rc[hinbr - 1] += level[lownbr] - level[hinbr];
          } else {
            // lca <- find(pleaf) with path halving.
            Idx last1 = pleaf, last2 = set[last1 - 1], lca = set[last2 - 1];
            while (lca != last2) {
              set[last1 - 1] = lca;
              last1 = lca;
              last2 = set[last1 - 1];
              lca = set[last2 - 1];
            }
            rc[hinbr - 1] += level[lownbr] - level[lca];
            --weight[lca];
          }
          prvlf[hinbr - 1] = lownbr;
          lflag = 1;
        }
        prvnbr[hinbr - 1] = lownbr;
      }
    }
    Idx parent = et.PostParent(lownbr - 1);
    --weight[parent];
    if (lflag == 1 || nchild[lownbr] >= 2) xsup = lownbr;
    set[xsup - 1] = parent;
  }
  for (Int k = 1; k <= size; ++k) {
    Int temp = cc[k - 1] + weight[k];
    cc[k - 1] = temp;
    Int parent = et.PostParent(k - 1);
    if (parent != 0) cc[parent - 1] += temp;
  }
}

// findSupernodes (impl:42-80): fundamental supernodes, column i joins i-1
// when parent(i-1)==i and cc(i-1)==cc(i)+1, up to maxSize columns.
void find_supernodes(const ETree& et, const std::vector<Int>& cc, Int maxSize,
                     std::vector<Int>& supMembership, std::vector<Int>& xsuper) {
  Int size = (Int)cc.size();
  supMembership.assign(size, 0);
  Int nsuper = 1, supsize = 1;
  if (size == 0) {
    xsuper.assign(1, 1);
    return;
  }
  supMembership[0] = 1;
  for (Int i = 2; i <= size; i++) {
    Int prev_parent = et.PostParent(i - 2);
    if (prev_parent == i && cc[i - 2] == cc[i - 1] + 1) {
      if (supsize < maxSize || maxSize == 0) {
        ++supsize;
        supMembership[i - 1] = nsuper;
        continue;
      }
    }
    nsuper++;
    supsize = 1;
    supMembership[i - 1] = nsuper;
  }
  xsuper.assign(nsuper + 1, 0);
  Int lstsup = nsuper + 1;
  for (Int i = size; i >= 1; --i) {
    Int ksup = supMembership[i - 1];
    if (ksup != lstsup) xsuper[lstsup - 1] = i + 1;
    lstsup = ksup;
  }
  xsuper[0] = 1;
}

namespace {
// DisjointSet (include/sympack/ETree.hpp:12-99), including its find().
struct DisjointSet {
  std::vector<Int> pp, root;
  void init(Int n) {
    pp.assign(n, 0);
    root.assign(n, 0);
  }
  Int makeSet(Int i) {
    pp[i - 1] = i;
    return i;
  }
  Int link(Int s, Int t) {
    pp[s - 1] = t;
    return t;
  }
  Int find(Int i) {
    Int p = pp[i - 1], gp = pp[p - 1];
    while (gp != p) {
      i = makeSet(gp);
      p = pp[i - 1];
      gp = pp[p - 1];
    }
    return p;
  }
  void unite(Int s, Int t, Int r) {
    Int tSet = find(t);
    Int sSet = find(s);
    sSet = link(sSet, tSet);
    root[sSet - 1] = r;
  }
};
}  // namespace

// relaxSupernodes (impl:473-609): merge a supernode into its parent set when
// that set starts right after it and the explicit-zero fraction is below the
// nrelax/zrelax thresholds; then rewrite xsuper, supMembership and cc.
void relax_supernodes(const ETree& et, std::vector<Int>& cc, std::vector<Int>& supMembership,
                      std::vector<Int>& xsuper, const RelaxParams& params) {
  Int nsuper = (Int)xsuper.size() - 1;
  DisjointSet sets;
  sets.init(nsuper);
  std::vector<Int> ncols(nsuper), zeros(nsuper), newCC(nsuper);
  for (Int ksup = nsuper; ksup >= 1; --ksup) {
    Int cset = sets.makeSet(ksup);
    sets.root[cset - 1] = ksup;
    Int fstcol = xsuper[ksup - 1], lstcol = xsuper[ksup] - 1;
    ncols[ksup - 1] = lstcol - fstcol + 1;
    zeros[ksup - 1] = 0;
    newCC[ksup - 1] = cc[fstcol - 1];
  }
  for (Int ksup = nsuper; ksup >= 1; --ksup) {
    Int fstcol = xsuper[ksup - 1], lstcol = xsuper[ksup] - 1;
    Int width = ncols[ksup - 1];
    Int parent_col = et.PostParent(lstcol - 1);
    if (parent_col == 0) continue;
    Int parent_snode = supMembership[parent_col - 1];
    Int pset = sets.find(parent_snode);
    parent_snode = sets.root[pset - 1];
    bool merge = (parent_snode == ksup + 1);
    if (!merge) continue;
    Int parent_width = ncols[parent_snode - 1];
    Int parent_fc = xsuper[parent_snode - 1];
    Int totzeros = zeros[parent_snode - 1];
    Int merged_snode_size = width + parent_width;
    merge = false;
    if ((merged_snode_size <= params.maxSize || params.maxSize == 0) && params.nrelax0 > 0) {
      if (merged_snode_size <= params.nrelax0) {
        merge = true;
      } else if (merged_snode_size <= params.maxSize || params.maxSize == 0) {
        double nnzchild = cc[fstcol - 1];
        double nnzparent = cc[parent_fc - 1];
        double xnewzeros = width * (nnzparent + width - nnzchild);
        if (xnewzeros == 0) {
          merge = true;
        } else {
          double xtotzeros = (double)totzeros + xnewzeros;
          double xmerged = (double)merged_snode_size;
          double xtotsize = (xmerged * (xmerged + 1) / 2) + xmerged * (nnzparent - parent_width);
          double z = xtotzeros / xtotsize;
          totzeros += (Int)xnewzeros;
          double ptr_max = (double)std::numeric_limits<uint64_t>::max();
          if (xtotsize * sizeof(double) < ptr_max) {
            if (merged_snode_size <= params.nrelax1 && z < params.zrelax0)
              merge = true;
            else if (merged_snode_size <= params.nrelax2 && z < params.zrelax1)
              merge = true;
            else if (z < params.zrelax2)
              merge = true;
          }
        }
      }
    }
    if (merge) {
      ncols[ksup - 1] += ncols[parent_snode - 1];
      zeros[ksup - 1] = totzeros;
      newCC[ksup - 1] = width + newCC[parent_snode - 1];
      sets.unite(ksup, parent_snode, ksup);
    }
  }
  std::vector<Int> relXSuper(nsuper + 1);
  Int nrSuper = 0;
  for (Int ksup = 1; ksup <= nsuper; ++ksup) {
    Int kset = sets.find(ksup);
    if (ksup == sets.root[kset - 1]) {
      relXSuper[nrSuper] = xsuper[ksup - 1];
      newCC[nrSuper] = newCC[ksup - 1];
      ++nrSuper;
    }
  }
  relXSuper[nrSuper] = xsuper[nsuper];
  relXSuper.resize(nrSuper + 1);
  for (Int ksup = 1; ksup <= nrSuper; ++ksup) {
    Int fstcol = relXSuper[ksup - 1], lstcol = relXSuper[ksup] - 1;
    for (Int col = fstcol; col <= lstcol; ++col) {
      supMembership[col - 1] = ksup;
      cc[col - 1] = newCC[ksup - 1] - (col - fstcol);
    }
  }
  xsuper = relXSuper;
}

// symbolicFactorizationRelaxedDist (impl:613-938) on one rank: the row
// structure of every supernode is the union of its children's structures
// (beyond their own columns), its own columns, and the permuted matrix
// entries below its first column (of its first column only when relaxation
// is off).  The reference builds the union in a sorted linked list, inserting
// each new row by a walk from the head, which is quadratic in the width of
// wide separators; the same sorted set is formed here with a marker array
// and linear merges of the (sorted) child structures.  The size check
// against the column count is the reference's.
void symbolic_structure(const Graph& expanded, const Ordering& ord, const ETree& et,
                        const std::vector<Int>& cc, const std::vector<Int>& xsuper,
                        const std::vector<Int>& supMembership, bool first_col_only,
                        std::vector<Ptr>& xlindx, std::vector<Idx>& lindx) {
  Int size = expanded.n;
  Int nsuper = (Int)xsuper.size() - 1;
  // children of every supernode, in the order the reference's merge list
  // holds them (the latest linked first)
  std::vector<Int> mrglnk(nsuper, 0);
  std::vector<Int> marker(size, 0);
  xlindx.assign(nsuper + 1, 0);
  Ptr point = 1;
  for (Int ksup = 1; ksup <= nsuper; ++ksup) {
    xlindx[ksup - 1] = point;
    point += cc[xsuper[ksup - 1] - 1];
  }
  xlindx[nsuper] = point;
  lindx.assign(point - 1, 0);
  // rows of the supernode past its own columns, kept sorted: children's
  // structures are sorted, so each child's new rows are merged in linear
  // time; the matrix entries are sorted on their own and merged last
  std::vector<Idx> tail, tmp, merged;
  for (Int ksup = 1; ksup <= nsuper; ++ksup) {
    const Int fstcol = xsuper[ksup - 1], lstcol = xsuper[ksup] - 1;
    const Int width = lstcol - fstcol + 1;
    const Int length = cc[fstcol - 1];
    for (Int row = fstcol; row <= lstcol; ++row) marker[row - 1] = ksup;
    tail.clear();
    auto merge_tmp = [&]() {
      if (tmp.empty()) return;
      if (tail.empty()) {
        tail.swap(tmp);
      } else {
        merged.resize(tail.size() + tmp.size());
        std::merge(tail.begin(), tail.end(), tmp.begin(), tmp.end(), merged.begin());
        tail.swap(merged);
      }
      tmp.clear();
    };
    for (Int jsup = mrglnk[ksup - 1]; jsup != 0 && width + (Int)tail.size() < length;
         jsup = mrglnk[jsup - 1]) {
      const Int jwidth = xsuper[jsup] - xsuper[jsup - 1];
      for (Ptr jptr = xlindx[jsup - 1] + jwidth; jptr <= xlindx[jsup] - 1; ++jptr) {
        const Idx r = lindx[jptr - 1];
        if (marker[r - 1] != ksup) {
          marker[r - 1] = ksup;
          tmp.push_back(r);
        }
      }
      merge_tmp();
    }
    if (width + (Int)tail.size() < length) {
      // column col of the permuted matrix is column perm[col] of A, its rows
      // mapped through invp (the order of the rows does not matter here)
      for (Int col = fstcol; col <= lstcol; ++col) {
        const Int old = ord.perm[col - 1];
        for (Ptr kptr = expanded.colptr[old - 1]; kptr <= expanded.colptr[old] - 1; ++kptr) {
          const Idx newi = ord.invp[expanded.rowind[kptr - 1] - 1];
          if (newi > (Idx)fstcol && marker[newi - 1] != ksup) {
            marker[newi - 1] = ksup;
            tmp.push_back(newi);
          }
        }
        if (first_col_only) break;
      }
      std::sort(tmp.begin(), tmp.end());
      merge_tmp();
    }
    if (width + (Int)tail.size() != length)
      throw std::logic_error("symbolic factorization: structure size " + std::to_string(width + tail.size()) +
                             " != column count " + std::to_string(length) + " for supernode " +
                             std::to_string(ksup));
    Idx* out = lindx.data() + (xlindx[ksup - 1] - 1);
    for (Int c = 0; c < width; ++c) out[c] = fstcol + c;
    std::copy(tail.begin(), tail.end(), out + width);
    if (length > width) {
      // linked into the merge list of the supernode of its parent column
      const Idx pcol = et.PostParent(fstcol + width - 1 - 1);
      const Int psup = supMembership[pcol - 1];
      mrglnk[ksup - 1] = mrglnk[psup - 1];
      mrglnk[psup - 1] = ksup;
    }
  }
}

void symbolic_factorization(Int n, const Ptr* colptr, const Idx* rowind, int baseval, int ordering,
                            const RelaxParams& relax, const Int* user_perm, Symbolic& out,
                            int refine) {
  if (n <= 0) throw std::invalid_argument("matrix dimension must be positive");
  out = Symbolic();
  out.n = n;
  Graph graph = expand_symmetric(n, colptr, rowind, baseval);
  Ordering& ord = out.ord;
  ETree& et = out.etree;
  std::vector<Int>& cc = out.cc;
  bool have_sgraph = false;
  Graph sgraph;
  switch (ordering) {
    case ORD_MMD:
    case ORD_RCM:
    case ORD_AMD:
      sgraph = graph;
      set_keep_diag(sgraph, false);
      if (ordering == ORD_MMD) order_mmd(sgraph, ord);
      if (ordering == ORD_RCM) order_rcm(sgraph, ord);
      if (ordering == ORD_AMD) order_amd(sgraph, ord);
      have_sgraph = true;
      break;
    case ORD_USER: {
      if (!user_perm)
        throw std::logic_error("When using USER, symPACKOptions.perm must be provided.\n");
      Int base = *std::min_element(user_perm, user_perm + n);
      ord.perm.resize(n);
      for (Int i = 0; i < n; i++) ord.perm[i] = user_perm[i] - base + 1;
      // The reference stores a 0-based invp here and then permutes with
      // base 1 (symPACKMatrix2D.hpp:2825-2834); a consistent 1-based invp is
      // built instead so that a user ordering is usable.
      ord.invp.assign(n, 0);
      for (Int i = 0; i < n; i++) {
        if (ord.perm[i] < 1 || ord.perm[i] > n || ord.invp[ord.perm[i] - 1] != 0)
          throw std::logic_error("USER ordering: perm is not a permutation");
        ord.invp[ord.perm[i] - 1] = i + 1;
      }
      break;
    }
    case ORD_NDBOX:
      order_ndbox(n, ord);
      break;
    case ORD_NDGRID:
      order_ndgrid(n, ord);
      break;
    case ORD_NATURAL:
      ord.perm.resize(n);
      std::iota(ord.perm.begin(), ord.perm.end(), 1);
      ord.invp = ord.perm;
      break;
    default:
      throw std::logic_error(
          "This ordering method is not supported by symPACK. Valid options are:NATURAL MMD AMD "
          "RCM NDBOX NDGRID USER \n");
  }

  if (!have_sgraph) {
    // symPACKMatrix2D.hpp:2901-2937: etree on the permuted expanded graph.
    // the elimination tree of the permuted matrix, which is unique: Liu's
    // algorithm on the unpermuted graph with rows mapped through invp
    // (construct_serial) finds the same tree without building the permuted
    // graph
    et.construct_serial(graph, ord);
    et.postorder(ord);
    // getLColRowCount on the graph permuted by the postordered ordering.
    lcol_row_count(graph, et, ord, cc, out.rc);
    if (ordering != ORD_NATURAL) et.sort_children(cc, ord);
  } else {
    // symPACKMatrix2D.hpp:2938-2970: diagonal re-inserted, serial etree.
    set_keep_diag(sgraph, true);
    et.construct_serial(sgraph, ord);
    et.postorder(ord);
    lcol_row_count(sgraph, et, ord, cc, out.rc);
    if (ordering != ORD_NATURAL) et.sort_children(cc, ord);
  }

  // statistics (symPACKMatrix2D.hpp:3008-3018)
  out.flops = 0.0;
  out.nnzL = 0;
  for (Int i = 0; i < n; ++i) {
    out.flops += std::pow((double)cc[i], 2.0);
    out.nnzL += cc[i];
  }
  out.cc_unrelaxed = cc;

  find_supernodes(et, cc, relax.maxSize, out.supMembership, out.xsuper);
  if (relax.nrelax0 > 0) relax_supernodes(et, cc, out.supMembership, out.xsuper, relax);

  bool first_col_only = (relax.nrelax0 == 0);
  symbolic_structure(graph, ord, et, cc, out.xsuper, out.supMembership, first_col_only, out.xlindx,
                     out.lindx);
  if (refine > 0) {
    // symPACKMatrix2D.hpp:3090-3111: reorder inside the supernodes, compose
    // (Ordering::Compose, src/Ordering.cpp:1301-1314), and build the row
    // structure again with the same partition (3246-3254)
    const std::vector<Int> invp2 =
        (refine == kRefineTspb || refine == kRefineTspbPlain)
            ? refine_supernodes_tsp(refine == kRefineTspb, n, graph, et, ord, out.xsuper, out.supMembership,
                                    out.xlindx, out.lindx)
            : refine_supernodes(refine / 10, refine % 10, n, out.xsuper, out.supMembership, out.xlindx,
                                out.lindx);
    for (Int i = 0; i < n; ++i) ord.invp[i] = invp2[ord.invp[i] - 1];
    for (Int i = 0; i < n; ++i) ord.perm[ord.invp[i] - 1] = i + 1;
    symbolic_structure(graph, ord, et, cc, out.xsuper, out.supMembership, first_col_only,
                       out.xlindx, out.lindx);
  }

  // supernodal etree (ETree::ToSupernodalETree, src/ETree.cpp:400-426)
  Int ns = out.nsuper();
  out.sparent.assign(ns, 0);
  for (Int s = 1; s <= ns; ++s) {
    Int lc = out.xsuper[s] - 1;
    Int pc = et.PostParent(lc - 1);
    out.sparent[s - 1] = pc == 0 ? 0 : out.supMembership[pc - 1];
  }
}

void build_cell_map(const Symbolic& s, CellMap& cm) {
  cm = CellMap();
  Int nsuper = s.nsuper();
  for (Int I = nsuper; I >= 1; --I) {
    Ptr lfi = s.xlindx[I - 1], lli = s.xlindx[I] - 1;
    Int Kprev = -1;
    Idx start = 0, prevRow = 0;
    Int contig = 0;
    bool open = false;
    auto close_block = [&]() {
      cm.block_first_row.push_back(start);
      cm.block_nrows.push_back(contig);
      cm.cell_nblocks.back()++;
    };
    for (Ptr p = lfi; p <= lli; ++p) {
      Idx row = s.lindx[p - 1];
      Int K = s.supMembership[row - 1];
      if (K != Kprev && K >= I) {
        if (open) close_block();
        cm.cell_K.push_back(K);
        cm.cell_I.push_back(I);
        cm.cell_nblocks.push_back(0);
        cm.cell_nrows.push_back(0);
        cm.cell_first_block.push_back((Ptr)cm.block_first_row.size());
        open = true;
        contig = 0;
        start = row;
        prevRow = row;
      }
      Kprev = K;
      if (row == prevRow + 1 || row == start) {
        ++contig;
      } else {
        close_block();
        contig = 1;
        start = row;
      }
      prevRow = row;
      cm.cell_nrows.back()++;
    }
    if (open) close_block();
  }
}

}  // namespace spk
