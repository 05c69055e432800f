// The symPACKMatrix2D adapter (include/sympack_amd.hpp) driven the way
// driver/run_sympack2D.cpp:74-128 drives symPACK's class, on the 64 x 64
// 5-point Laplacian (n = 4096, lower triangle, 1-based, MMD ordering).
//
// The option and matrix types below carry the member names of symPACK's
// symPACKOptions / RelaxationParameters (include/sympack/Types.hpp:20-188)
// and DistSparseMatrix / DistSparseMatrixGraph
// (include/sympack/DistSparseMatrix.hpp:46-73, DistSparseMatrixGraph.hpp
// :54-70), which the adapter's templates read.
//
// Modes:
//   symbolic   Init + SymbolicFactorization only (no GPU): accessor checks
//              and the logic_error paths
//   full       the whole call sequence on the GPU, LL^T and LDL^T, then the
//              "Matrix is not HPD." path
// Prints one line per check and "adapter_c1: OK" at the end; exit 0 on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "sympack_amd.hpp"

namespace ref_types {
typedef int64_t Int;
typedef uint64_t Ptr;
typedef uint32_t Idx;
enum class DecompositionType { LL, LDL };
struct RelaxationParameters {
  Int nrelax0 = 8, nrelax1 = 32, nrelax2 = 64, maxSize = 0;
  double zrelax0 = 0.8, zrelax1 = 0.1, zrelax2 = 0.05;
};
struct symPACKOptions {
  DecompositionType decomposition = DecompositionType::LL;
  std::string orderingStr = "MMD";
  std::string order_refinement_str = "NO";
  RelaxationParameters relax;
  int verbose = 0;
  int* perm = nullptr;
};
struct DistSparseMatrixGraph {
  Idx size = 0;
  Ptr nnz = 0;
  std::vector<Ptr> colptr;
  std::vector<Idx> rowind;
  int baseval = 1;
};
template <typename F>
struct DistSparseMatrix {
  DistSparseMatrixGraph Localg_;
  Int size = 0;
  Ptr nnz = 0;
  std::vector<F> nzvalLocal;
  DistSparseMatrixGraph& GetLocalGraph() { return Localg_; }
};
}  // namespace ref_types

using namespace ref_types;
typedef sympack_amd::symPACKMatrix2D<Ptr, Idx, double> Mat2D;

static int failures = 0;
#define CHECK(cond, ...)                    \
  do {                                      \
    if (cond) {                             \
      std::printf("ok   " __VA_ARGS__);     \
    } else {                                \
      std::printf("FAIL " __VA_ARGS__);     \
      ++failures;                           \
    }                                       \
    std::printf("\n");                      \
    std::fflush(stdout);                    \
  } while (0)

// 5-point Laplacian on a k x k grid, natural numbering i + k j, Dirichlet
// boundary: lower triangle incl. diagonal, 1-based, rows sorted.
static DistSparseMatrix<double> laplacian(int k, double diag) {
  DistSparseMatrix<double> A;
  const int n = k * k;
  A.size = n;
  auto& g = A.Localg_;
  g.size = n;
  g.colptr.push_back(1);
  for (int c = 0; c < n; ++c) {
    const int i = c % k, j = c / k;
    g.rowind.push_back(c + 1);
    A.nzvalLocal.push_back(diag);
    if (i + 1 < k) { g.rowind.push_back(c + 2); A.nzvalLocal.push_back(-1.0); }
    if (j + 1 < k) { g.rowind.push_back(c + k + 1); A.nzvalLocal.push_back(-1.0); }
    g.colptr.push_back(g.rowind.size() + 1);
  }
  g.nnz = A.nnz = g.rowind.size();
  return A;
}

// y = A x for the symmetric matrix held as its lower triangle (1-based)
static void matvec(DistSparseMatrix<double>& A, const double* x, double* y) {
  const int n = (int)A.size;
  auto& g = A.Localg_;
  for (int i = 0; i < n; ++i) y[i] = 0.0;
  for (int c = 0; c < n; ++c)
    for (Ptr q = g.colptr[c] - 1; q < g.colptr[c + 1] - 1; ++q) {
      const int r = (int)g.rowind[q] - 1;
      y[r] += A.nzvalLocal[q] * x[c];
      if (r != c) y[c] += A.nzvalLocal[q] * x[r];
    }
}

static double solve_and_check(const char* what, DistSparseMatrix<double>& A, const std::string& fact,
                              int nrhs) {
  symPACKOptions o;
  o.orderingStr = "MMD";
  o.decomposition = fact == "LDL" ? DecompositionType::LDL : DecompositionType::LL;
  Mat2D M;
  M.Init(o);
  M.SymbolicFactorization(A);
  M.DistributeMatrix(A);
  M.Factorize();
  const int n = (int)A.size;
  // x_true uniform in [-1, 1] from a fixed 64-bit LCG
  std::vector<double> xt((size_t)n * nrhs), b((size_t)n * nrhs), x((size_t)n * nrhs);
  uint64_t s = 0x9e3779b97f4a7c15ull;
  for (auto& v : xt) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    v = 2.0 * ((double)(s >> 11) / 9007199254740992.0) - 1.0;
  }
  for (int j = 0; j < nrhs; ++j) matvec(A, &xt[(size_t)j * n], &b[(size_t)j * n]);
  x = b;
  M.Solve(x.data(), nrhs, (int)x.size());
  M.GetSolution(x.data(), nrhs);
  // check_solution (driver/utils.hpp:95-142): ||AX - B||_F / ||B||_F
  std::vector<double> ax(n);
  double rn = 0, bn = 0, en = 0, xn = 0;
  for (int j = 0; j < nrhs; ++j) {
    matvec(A, &x[(size_t)j * n], ax.data());
    for (int i = 0; i < n; ++i) {
      const double r = ax[i] - b[(size_t)j * n + i];
      rn += r * r;
      bn += b[(size_t)j * n + i] * b[(size_t)j * n + i];
      const double e = x[(size_t)j * n + i] - xt[(size_t)j * n + i];
      en += e * e;
      xn += xt[(size_t)j * n + i] * xt[(size_t)j * n + i];
    }
  }
  const double rel = std::sqrt(rn / bn), fwd = std::sqrt(en / xn);
  CHECK(rel < 1e-13 && fwd < 1e-11, "%s: relative residual %.3e, forward error %.3e (nsuper %lld)", what, rel,
        fwd, (long long)M.TotalSupernodeCnt());
  return rel;
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "full";
  int ac = 0;
  char** av = nullptr;
  symPACK_Init(&ac, &av);
  DistSparseMatrix<double> A = laplacian(64, 4.0);
  CHECK(A.size == 4096 && A.nnz == 12160, "C1 matrix: n = %lld, stored lower entries = %llu", (long long)A.size,
        (unsigned long long)A.nnz);

  {  // symbolic analysis and accessors
    symPACKOptions o;
    Mat2D M;
    M.Init(o);
    M.SymbolicFactorization(A);
    const auto xs = M.GetSupernodalPartition();
    const int ns = (int)M.TotalSupernodeCnt();
    CHECK((int)xs.size() == ns + 1 && xs.front() == 1 && xs.back() == 4097,
          "partition: %d supernodes covering 1..4096", ns);
    const auto ord = M.GetOrdering();
    std::vector<int> seen(4097, 0);
    bool perm_ok = ord.perm.size() == 4096;
    for (size_t i = 0; perm_ok && i < ord.perm.size(); ++i) {
      const int p = ord.perm[i];
      perm_ok = p >= 1 && p <= 4096 && !seen[p] && ord.invp[p - 1] == (int)i + 1;
      if (perm_ok) seen[p] = 1;
    }
    CHECK(perm_ok, "ordering: perm is a permutation and invp its inverse");
    const auto sup = M.GetSupMembership();
    bool sup_ok = (int)sup.size() == 4096;
    for (int I = 0; sup_ok && I < ns; ++I)
      for (int c = xs[I]; c < xs[I + 1]; ++c) sup_ok = sup_ok && sup[c - 1] == I + 1;
    CHECK(sup_ok, "supernode membership follows the partition");
    const auto par = M.GetETree();
    bool tree_ok = true;
    for (int v = 1; v <= 4096; ++v) tree_ok = tree_ok && (par[v - 1] == 0 || par[v - 1] > v);
    CHECK(tree_ok, "elimination tree is postordered");
  }
  {  // logic_error paths
    symPACKOptions o;
    o.orderingStr = "METIS";
    Mat2D M;
    bool threw = false;
    try {
      M.Init(o);
    } catch (const std::logic_error& e) {
      threw = std::strstr(e.what(), "not supported") != nullptr;
    }
    CHECK(threw, "unknown ordering name raises logic_error");
    o.orderingStr = "USER";
    threw = false;
    try {
      M.Init(o);
    } catch (const std::logic_error&) {
      threw = true;
    }
    CHECK(threw, "USER ordering without perm raises logic_error");
  }

  if (mode == "full") {
    solve_and_check("LL^T MMD nrhs=1", A, "LL", 1);
    solve_and_check("LL^T MMD nrhs=7", A, "LL", 7);
    solve_and_check("LDL^T MMD nrhs=1", A, "LDL", 1);
    DistSparseMatrix<double> B = laplacian(64, 4.0);
    B.nzvalLocal[B.Localg_.colptr[100] - 1] = -4.0;  // diagonal of column 101
    symPACKOptions o;
    Mat2D M;
    M.Init(o);
    M.SymbolicFactorization(B);
    M.DistributeMatrix(B);
    bool hpd_error = false;
    try {
      M.Factorize();
    } catch (const std::runtime_error& e) {
      hpd_error = std::string(e.what()) == "Matrix is not HPD.";
    }
    CHECK(hpd_error, "indefinite matrix: Factorize raises runtime_error(\"Matrix is not HPD.\")");
  }
  symPACK_Finalize();
  std::printf("adapter_c1: %s\n", failures ? "FAILED" : "OK");
  return failures ? 1 : 0;
}
