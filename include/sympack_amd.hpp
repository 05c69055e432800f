/*
 * symPACKMatrix2D over the C ABI of sympack_amd.h: one rank, one GPU.
 *
 * Keeps the reference class's name, template parameters, method names,
 * argument meaning, call order and exceptions
 * (include/sympack/symPACKMatrix2D.hpp:2469-2631, driver/run_sympack2D.cpp
 * :74-128), so code written against symPACK's 2D class compiles against this
 * header with the include and the namespace changed.
 *
 * The methods that take symPACK's option and matrix objects are templates
 * over those types and read only members the reference declares:
 *   options (symPACKOptions, include/sympack/Types.hpp:105-188):
 *     orderingStr, decomposition, relax.{nrelax0,nrelax1,nrelax2,maxSize,
 *     zrelax0,zrelax1,zrelax2}, perm, order_refinement_str, verbose
 *   matrix (DistSparseMatrix<T>, include/sympack/DistSparseMatrix.hpp:46-73):
 *     size, nzvalLocal, GetLocalGraph().{colptr,rowind,baseval}
 * so symPACK's own classes work, and so does any type with those members.
 *
 * Exceptions, as the reference raises them: std::logic_error for a bad
 * ordering name or a USER ordering without perm (symPACKMatrix2D.hpp:
 * 2732-2790); std::runtime_error("Matrix is not HPD.") when a pivot fails
 * (src/lapack.cpp:293-306; the reference then hangs, this returns);
 * std::bad_alloc when device memory runs out; std::runtime_error for any
 * other device error.
 */
#ifndef SYMPACK_AMD_HPP_
#define SYMPACK_AMD_HPP_

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sympack_amd.h"

namespace sympack_amd {

namespace detail {
inline int check(int rc) {
  if (rc == -2) throw std::logic_error(spk_last_error());
  if (rc == -3) throw std::bad_alloc();
  if (rc < 0) throw std::runtime_error(spk_last_error());
  return rc;
}

/* options.orderingStr -> ordering code (symPACKMatrix2D.hpp:2732-2790).
 * NDBOX and NDGRID (src/Ordering.cpp:1289-1300) are accepted as well. */
inline int ordering_code(const std::string& s) {
  if (s == "MMD") return SPK_ORDER_MMD;
  if (s == "RCM") return SPK_ORDER_RCM;
  if (s == "AMD") return SPK_ORDER_AMD;
  if (s == "NATURAL") return SPK_ORDER_NATURAL;
  if (s == "NDBOX") return SPK_ORDER_NDBOX;
  if (s == "NDGRID") return SPK_ORDER_NDGRID;
  if (s == "USER") return SPK_ORDER_USER;
  throw std::logic_error(
      "This ordering method is not supported by symPACK. Valid options are:"
      "NATURAL MMD AMD RCM NDBOX NDGRID USER \n");
}

/* options.order_refinement_str -> spk_options.refine
 * (symPACKMatrix2D.hpp:3090-3263): SET is SET31; TSPB is the tour with
 * skeleton compression (40), any other name starting with TSPB the tour
 * without it (41); names the reference does not act on leave the ordering
 * unchanged. */
inline int refine_code(const std::string& r) {
  if (r == "TSPB") return 40;
  if (r.compare(0, 4, "TSPB") == 0) return 41;
  if (r == "SET" || r == "SET31") return 31;
  if (r == "SET10") return 10;
  if (r == "SET11") return 11;
  if (r == "SET20") return 20;
  if (r == "SET21") return 21;
  if (r == "SET30") return 30;
  return 0;
}

template <class D>
inline int decomposition_code(const D& d) {
  return static_cast<int>(d) == 0 ? SPK_FACT_LL : SPK_FACT_LDL;  // DecompositionType {LL, LDL}
}
}  // namespace detail

/* Ordering as symPACK::Ordering exposes it: perm[new-1] = old,
 * invp[old-1] = new, 1-based. */
template <typename int_t>
struct OrderingView {
  std::vector<int_t> perm, invp;
};

template <typename colptr_t, typename rowind_t, typename T, typename int_t = int>
class symPACKMatrix2D {
  static_assert(std::is_same<T, double>::value, "sympack_amd factors double-precision matrices");

 public:
  symPACKMatrix2D() { detail::check(spk_create(&h_)); }
  ~symPACKMatrix2D() {
    if (h_) spk_destroy(h_);
  }
  symPACKMatrix2D(const symPACKMatrix2D&) = delete;
  symPACKMatrix2D& operator=(const symPACKMatrix2D&) = delete;

  /* symPACKMatrix2D::Init (symPACKMatrix2D.hpp:2649-2695) */
  template <class Options>
  void Init(Options& options) {
    spk_options so;
    spk_default_options(&so);
    so.ordering = detail::ordering_code(options.orderingStr);
    if (so.ordering == SPK_ORDER_USER && options.perm == nullptr)
      throw std::logic_error("When using USER, symPACKOptions.perm must be provided.\n");
    so.decomposition = detail::decomposition_code(options.decomposition);
    so.nrelax0 = (int)options.relax.nrelax0;
    so.nrelax1 = (int)options.relax.nrelax1;
    so.nrelax2 = (int)options.relax.nrelax2;
    so.maxsize = (int)options.relax.maxSize;
    so.zrelax0 = options.relax.zrelax0;
    so.zrelax1 = options.relax.zrelax1;
    so.zrelax2 = options.relax.zrelax2;
    so.perm = options.perm;
    so.refine = detail::refine_code(options.order_refinement_str);
    so.verbose = options.verbose;
    detail::check(spk_init(h_, &so));
  }

  /* symPACKMatrix2D::SymbolicFactorization (symPACKMatrix2D.hpp:2699-6334):
   * the lower-triangular CSC graph of the matrix. */
  template <class Matrix>
  void SymbolicFactorization(Matrix& pMat) {
    auto& g = pMat.GetLocalGraph();
    const int32_t n = (int32_t)pMat.size;
    std::vector<int64_t> cp(g.colptr.begin(), g.colptr.end());
    std::vector<int32_t> ri(g.rowind.begin(), g.rowind.end());
    if ((int64_t)cp.size() != (int64_t)n + 1) throw std::logic_error("colptr must hold size + 1 entries");
    detail::check(spk_symbolic_factorization(h_, n, cp.data(), ri.data(), g.baseval));
    n_ = n;
  }

  /* symPACKMatrix2D::DistributeMatrix (symPACKMatrix2D.hpp:6337-6576) */
  template <class Matrix>
  void DistributeMatrix(Matrix& pMat) {
    detail::check(spk_distribute_matrix(h_, pMat.nzvalLocal.data()));
  }

  /* symPACKMatrix2D::Factorize (symPACKMatrix2D.hpp:6579-6639) */
  void Factorize() {
    if (detail::check(spk_factorize(h_)) > 0) throw std::runtime_error("Matrix is not HPD.");
  }

  /* symPACKMatrix2D::Solve (symPACKMatrix2D.hpp:6643-6710): RHS is read,
   * column-major n x nrhs; rhs_size and Xptr are unused, as in the 2D path. */
  void Solve(T* RHS, int nrhs, int rhs_size, T* Xptr = nullptr) {
    (void)Xptr;
    detail::check(spk_solve(h_, RHS, nrhs, rhs_size));
  }

  /* symPACKMatrix2D::GetSolution (symPACKMatrix2D.hpp:6713-6757) */
  void GetSolution(T* B, int nrhs) { detail::check(spk_get_solution(h_, B, nrhs)); }

  /* symPACKMatrixBase accessors (symPACKMatrixBase.hpp:56-103), 1-based */
  int_t TotalSupernodeCnt() {
    spk_stats st;
    detail::check(spk_get_stats(h_, &st));
    return (int_t)st.nsuper;
  }
  std::vector<int_t> GetSupernodalPartition() {
    std::vector<int32_t> x((size_t)TotalSupernodeCnt() + 1);
    detail::check(spk_get_supernodal_partition(h_, x.data()));
    return std::vector<int_t>(x.begin(), x.end());
  }
  std::vector<int_t> GetSupMembership() {
    std::vector<int32_t> s((size_t)n_);
    detail::check(spk_get_sup_membership(h_, s.data()));
    return std::vector<int_t>(s.begin(), s.end());
  }
  OrderingView<int_t> GetOrdering() {
    std::vector<int32_t> p((size_t)n_), q((size_t)n_);
    detail::check(spk_get_ordering(h_, p.data(), q.data()));
    OrderingView<int_t> o;
    o.perm.assign(p.begin(), p.end());
    o.invp.assign(q.begin(), q.end());
    return o;
  }
  /* postordered elimination tree, parent of every column (0 = root) */
  std::vector<int_t> GetETree() {
    std::vector<int32_t> p((size_t)n_);
    detail::check(spk_get_etree(h_, p.data()));
    return std::vector<int_t>(p.begin(), p.end());
  }

  /* the underlying C handle, for the entry points beyond the reference's */
  spk_handle handle() const { return h_; }

 private:
  spk_handle h_ = nullptr;
  int32_t n_ = 0;
};

}  // namespace sympack_amd

#endif /* SYMPACK_AMD_HPP_ */
