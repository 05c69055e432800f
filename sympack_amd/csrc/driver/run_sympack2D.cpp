// run_sympack2D-compatible driver (driver/run_sympack2D.cpp:22-140 of the
// reference): reads a Harwell-Boeing / Rutherford-Boeing, binary CSC or
// text "matrix" lower triangle (ReadMatrix, utility.hpp:1693-1716),
// builds b = A x_true with the reference's alternating +-1 x_true
// (driver/utils.hpp:9-91), runs Init / SymbolicFactorization /
// DistributeMatrix / Factorize / Solve / GetSolution on one GPU and prints the
// same timing lines and the relative residual (driver/utils.hpp:95-142).
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sympack_amd.h"

static double get_time() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void check(int rc, const char* what) {
  if (rc < 0) {
    std::fprintf(stderr, "%s failed: %s\n", what, spk_last_error());
    std::exit(1);
  }
}

int main(int argc, char** argv) {
  check(symPACK_Init(&argc, &argv), "symPACK_Init");
  std::string filename, informat = "HARWELL_BOEING";
  int nrhs = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "-in") && i + 1 < argc) filename = argv[i + 1];
    if (!std::strcmp(argv[i], "-inf") && i + 1 < argc) informat = argv[i + 1];
    if (!std::strcmp(argv[i], "-nrhs") && i + 1 < argc) nrhs = std::atoi(argv[i + 1]);
  }
  if (filename.empty()) {
    std::fprintf(stderr, "usage: %s -in <file> [-inf HARWELL_BOEING|RB|HB|CSC|matrix] [-ordering MMD|...] [-nrhs k]\n",
                 argv[0]);
    return 1;
  }
  if (informat != "HARWELL_BOEING" && informat != "RB" && informat != "HB" && informat != "CSC" &&
      informat != "matrix") {
    std::fprintf(stderr, "Unknown matrix format. (%s)\n", informat.c_str());
    return 1;
  }
  spk_options opt;
  spk_default_options(&opt);
  check(spk_options_from_args(&opt, argc - 1, (const char* const*)(argv + 1)), "options");

  int32_t n = 0;
  int64_t nnz = 0;
  check(spk_read_matrix(filename.c_str(), informat.c_str(), &n, &nnz), "ReadMatrix");
  std::vector<int64_t> colptr(n + 1);
  std::vector<int32_t> rowind(nnz);
  std::vector<double> val(nnz);
  check(spk_read_matrix_data(filename.c_str(), informat.c_str(), colptr.data(), rowind.data(), val.data()),
        "ReadMatrix");

  // b = A x_true with x_true alternating +1/-1 (utils.hpp:22-28)
  std::vector<double> xtrue((size_t)n * nrhs), rhs((size_t)n * nrhs, 0.0);
  double v = 1.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < nrhs; ++j) {
      xtrue[i + (size_t)j * n] = v;
      v = -v;
    }
  auto spmv = [&](const std::vector<double>& x, std::vector<double>& y) {
    std::fill(y.begin(), y.end(), 0.0);
    for (int k = 0; k < nrhs; ++k)
      for (int c = 0; c < n; ++c)
        for (int64_t p = colptr[c] - 1; p < colptr[c + 1] - 1; ++p) {
          int r = rowind[p] - 1;
          y[r + (size_t)k * n] += val[p] * x[c + (size_t)k * n];
          if (r > c) y[c + (size_t)k * n] += val[p] * x[r + (size_t)k * n];
        }
  };
  spmv(xtrue, rhs);

  const char* ordnames[] = {"NATURAL", "RCM", "MMD", "AMD", "NDBOX", "NDGRID"};
  std::printf("Ordering: %s\n", opt.ordering <= 5 ? ordnames[opt.ordering] : "USER");
  spk_handle h;
  check(spk_create(&h), "create");
  check(spk_init(h, &opt), "Init");
  double t0 = get_time();
  check(spk_symbolic_factorization(h, n, colptr.data(), rowind.data(), 1), "SymbolicFactorization");
  check(spk_distribute_matrix(h, val.data()), "DistributeMatrix");
  double t1 = get_time();
  std::printf("Initialization time: %g seconds\n", t1 - t0);
  spk_stats st;
  spk_get_stats(h, &st);
  if (opt.verbose) {
    std::printf("Flops: %g\n", st.flops);
    std::printf("NNZ in L factor: %lld\n", (long long)st.nnz_l);
  }
  t0 = get_time();
  int info = spk_factorize(h);
  t1 = get_time();
  if (info > 0) {
    std::fprintf(stderr, "Runtime error: Matrix is not HPD. (column %d)\n", info);
    return 2;
  }
  check(info, "Factorize");
  std::printf("Factorization time: %g seconds\n", t1 - t0);
  if (nrhs > 0) {
    std::printf("Starting solve 2D\n");
    std::vector<double> x((size_t)n * nrhs);
    t0 = get_time();
    check(spk_solve(h, rhs.data(), nrhs, (int64_t)x.size()), "Solve");
    t1 = get_time();
    std::printf("Solve 2D time: %g seconds\n", t1 - t0);
    check(spk_get_solution(h, x.data(), nrhs), "GetSolution");
    std::vector<double> ax((size_t)n * nrhs);
    spmv(x, ax);
    double nr = 0, nb = 0;
    for (size_t i = 0; i < ax.size(); ++i) {
      nr += (ax[i] - rhs[i]) * (ax[i] - rhs[i]);
      nb += rhs[i] * rhs[i];
    }
    std::printf("Norm of residual after factorization/solve is %g\n", std::sqrt(nr) / std::sqrt(nb));
  }
  spk_destroy(h);
  symPACK_Finalize();
  return 0;
}
