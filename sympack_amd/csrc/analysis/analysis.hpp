// Host-side symbolic analysis for the supernodal Cholesky / static LDL^T path.
//
// Everything in this header is a restatement of symPACK's analysis chain
// (ordering -> elimination tree -> postorder -> SortChildren -> column counts
// -> fundamental supernodes -> relaxation -> relaxed symbolic factorization ->
// 2D cell/block map).  The supernode partition and the row structure must be
// bit-identical to symPACK's, so integer tie-breaks and the double-precision
// relaxation thresholds follow the reference line by line.  Indices are
// 1-based inside these routines, as in the reference, unless noted.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace spk {

using Int = int32_t;   // symPACK Int  (include/sympack/datatypes.hpp:12)
using Idx = int32_t;   // symPACK Idx  (uint32 in the reference; values < 2^31 here)
using Ptr = int64_t;   // symPACK Ptr  (uint64 in the reference)

// Ordering kinds; numeric values follow the enum order of
// include/sympack/Types.hpp:101 (NATURAL,RCM,MMD,AMD,NDBOX,NDGRID,...,USER).
enum OrderingType : int {
  ORD_NATURAL = 0,
  ORD_RCM = 1,
  ORD_MMD = 2,
  ORD_AMD = 3,
  ORD_NDBOX = 4,
  ORD_NDGRID = 5,
  ORD_USER = 10,
};

// Symmetric adjacency graph in compressed-column form, 1-based
// (SparseMatrixGraph, include/sympack/DistSparseMatrixGraph.hpp).
struct Graph {
  Int n = 0;
  std::vector<Ptr> colptr;   // n+1 entries, 1-based positions
  std::vector<Idx> rowind;   // 1-based row indices
  bool keep_diag = true;
  Ptr nnz() const { return (Ptr)rowind.size(); }
};

// Relaxation parameters (include/sympack/Types.hpp:22-95).
struct RelaxParams {
  Int nrelax0 = 8, nrelax1 = 32, nrelax2 = 64, maxSize = 0;
  double zrelax0 = 0.8, zrelax1 = 0.1, zrelax2 = 0.05;
};

struct Ordering {
  std::vector<Int> perm;  // perm[new-1] = old (1-based)
  std::vector<Int> invp;  // invp[old-1] = new (1-based)
  // Ordering::Compose (src/Ordering.cpp:1301-1315)
  void compose(const std::vector<Int>& invp2);
  // Ordering::GetRelativeInvp (src/Ordering.cpp:1289-1299)
  void relative_invp(const std::vector<Int>& frominvp, std::vector<Int>& relinvp) const;
};

// ---- graph utilities (src/DistSparseMatrixGraph.cpp) ----------------------
// ExpandSymmetric on a lower-triangular pattern (cpp:929-1083): every
// strictly-lower entry (i,j) also appears as (j,i); columns end up sorted.
Graph expand_symmetric(Int n, const Ptr* colptr, const Idx* rowind, int baseval);
// SetKeepDiag (cpp:39-97).
void set_keep_diag(Graph& g, bool keep);
// Permute an expanded graph (cpp:491-640): new column invp[j] holds invp[rows].
Graph permute_graph(const Graph& g, const std::vector<Int>& invp);

// ---- orderings -------------------------------------------------------------
// Liu's multiple minimum degree, as called by Ordering::MMD
// (src/Ordering.cpp:103-218 -> ordmmd.F -> genmmd.F, mmdint/mmdelm/mmdupd/mmdnum).
// g: expanded, no diagonal, 1-based.
void order_mmd(const Graph& g, Ordering& ord);
// Reverse Cuthill-McKee (Ordering.cpp:221-330 -> genrcm/rcm/fnroot/rootls/degree).
void order_rcm(const Graph& g, Ordering& ord);
// Approximate minimum degree (Ordering.cpp:411-525 -> amdbar.F).
void order_amd(const Graph& g, Ordering& ord);
// Geometric nested dissection on a k^3 box / k^2 grid
// (Ordering.cpp:344-408 -> boxnd.f, gridnd.f).
void order_ndbox(Int n, Ordering& ord);
void order_ndgrid(Int n, Ordering& ord);
// Raw routines returning the inverse ordering (BOX/GRID arrays), 1-based.
int boxnd(Int p, Int q, Int r, std::vector<Int>& box);
int gridnd(Int p, Int q, std::vector<Int>& grid);

// ---- elimination tree (src/ETree.cpp) --------------------------------------
struct ETree {
  Int n = 0;
  bool postordered = false;
  std::vector<Int> parent;    // 1-based parent, 0 = root
  std::vector<Int> poparent;  // parent after postordering
  Int PostParent(Int i) const { return poparent[i]; }
  // ConstructETree(SparseMatrixGraph&, Ordering&) (cpp:317-393)
  void construct_serial(const Graph& sg, const Ordering& ord);
  // ConstructETree(DistSparseMatrixGraph&) on the permuted graph (cpp:227-314)
  void construct_permuted(const Graph& pg);
  // PostOrderTree (cpp:59-119)
  void postorder(Ordering& ord);
  // SortChildren (cpp:122-225)
  void sort_children(std::vector<Int>& cc, Ordering& ord);
};

// ---- symbolic factorization -------------------------------------------------
struct Symbolic {
  Int n = 0;
  Ordering ord;
  ETree etree;
  std::vector<Int> cc_unrelaxed;   // column counts before relaxation
  std::vector<Int> cc;             // column counts after relaxation
  std::vector<Int> rc;             // row counts
  std::vector<Int> xsuper;         // nsuper+1, 1-based first columns
  std::vector<Int> supMembership;  // n, 1-based supernode of each column
  std::vector<Ptr> xlindx;         // nsuper+1, 1-based positions into lindx
  std::vector<Idx> lindx;          // row structure of every supernode (1-based rows)
  std::vector<Int> sparent;        // supernodal etree parent (1-based, 0 = root)
  double flops = 0.0;              // sum cc_j^2 over unrelaxed counts
  int64_t nnzL = 0;                // sum cc_j over unrelaxed counts
  Int nsuper() const { return (Int)xsuper.size() - 1; }
};

constexpr int kRefineTspb = 40;       // TSPB: tour with skeleton compression
constexpr int kRefineTspbPlain = 41;  // any other TSPB* name: without compression

// Full analysis as in symPACKMatrix2D::SymbolicFactorization
// (include/sympack/symPACKMatrix2D.hpp:2699-3081).  The input is the
// caller's lower-triangular CSC pattern (incl. diagonal).
// user_perm: for ORD_USER only (any base; shifted by its minimum).
// refine: 0, or 10*ordflag + altflag of the supernode reordering
// (-refine SET10 .. SET31; symPACKMatrix2D.hpp:3090-3111), or
// kRefineTspb / kRefineTspbPlain (-refine TSPB / other TSPB* names,
// symPACKMatrix2D.hpp:3113-3263).
void symbolic_factorization(Int n, const Ptr* colptr, const Idx* rowind, int baseval,
                            int ordering, const RelaxParams& relax, const Int* user_perm,
                            Symbolic& out, int refine = 0);

// Traveling-salesman reordering inside the supernodes (tspb.cpp; -refine
// TSPB with skeleton compression, other TSPB* names without): new position
// of every column (1-based), to be composed into the ordering.  graph: the
// expanded pattern in the original numbering; et, ord: the postordered tree
// and the ordering the structure was built with.
std::vector<Int> refine_supernodes_tsp(bool compress, Int n, const Graph& graph, const ETree& et,
                                       const Ordering& ord, const std::vector<Int>& xsuper,
                                       const std::vector<Int>& snode, const std::vector<Ptr>& xlindx,
                                       const std::vector<Idx>& lindx);

// Reordering inside the supernodes (refine.cpp): new position of every
// column (1-based), to be composed into the ordering.
std::vector<Int> refine_supernodes(int ordflag, int altflag, Int n, const std::vector<Int>& xsuper,
                                   const std::vector<Int>& snode, const std::vector<Ptr>& xlindx,
                                   const std::vector<Idx>& lindx);

// Pieces, exposed for tests.
void lcol_row_count(const Graph& sg, const ETree& et, const Ordering& ord,
                    std::vector<Int>& cc, std::vector<Int>& rc);
void find_supernodes(const ETree& et, const std::vector<Int>& cc, Int maxSize,
                     std::vector<Int>& supMembership, std::vector<Int>& xsuper);
void relax_supernodes(const ETree& et, std::vector<Int>& cc, std::vector<Int>& supMembership,
                      std::vector<Int>& xsuper, const RelaxParams& params);
void symbolic_structure(const Graph& expanded, const Ordering& ord, const ETree& et,
                        const std::vector<Int>& cc, const std::vector<Int>& xsuper,
                        const std::vector<Int>& supMembership, bool first_col_only,
                        std::vector<Ptr>& xlindx, std::vector<Idx>& lindx);

// 2D cell/block map (symPACKMatrix2D.hpp:3317-3396): for every supernode I
// (descending), each distinct row-supernode K>=I of I's structure forms
// cell (K,I); blocks are maximal runs of consecutive rows.
struct CellMap {
  std::vector<Int> cell_K, cell_I;        // 1-based supernode pair per cell
  std::vector<Int> cell_nblocks, cell_nrows;
  std::vector<Ptr> cell_first_block;      // index into block arrays
  std::vector<Idx> block_first_row;       // 1-based first row
  std::vector<Int> block_nrows;
};
void build_cell_map(const Symbolic& s, CellMap& cm);

// Rutherford-Boeing / Harwell-Boeing reader mirroring ReadHB_PARA
// (include/sympack/utility.hpp:1274-1482) on one rank.  Returns a 1-based
// lower-triangular CSC with rows sorted per column.
struct CscMatrix {
  Int n = 0;
  std::vector<Ptr> colptr;
  std::vector<Idx> rowind;
  std::vector<double> nzval;
};
void read_rb(const std::string& filename, CscMatrix& out);
// Binary CSC (ParaReadDistSparseMatrix, utility.hpp:1037-1265) and the text
// "matrix" format (ReadDistSparseMatrixFormatted, utility.hpp:1486-1688), and
// ReadMatrix's dispatch by format name (utility.hpp:1702-1716): "CSC",
// "HB" / "RB" / "HARWELL_BOEING", "matrix"; other names raise logic_error.
void read_csc_binary(const std::string& filename, CscMatrix& out);
void write_csc_binary(const std::string& filename, const CscMatrix& a);
void read_matrix_formatted(const std::string& filename, CscMatrix& out);
void write_matrix_formatted(const std::string& filename, const CscMatrix& a);
void read_matrix(const std::string& filename, const std::string& format, CscMatrix& out);
void write_rb(const std::string& filename, const CscMatrix& a, const std::string& title,
              const std::string& key);

}  // namespace spk
