/*
 * Oracle: CPU restatement of symPACK's supernodal numeric factorization and
 * triangular solve (2D path, one rank).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load this library, and only as the checker;
 * the product path (libsympack_amd.so) never links or calls it.
 *
 * Storage follows blockCell_t (include/sympack/symPACKMatrix2D.hpp:499-943):
 * supernode I owns the rows of its structure x its w_I columns, row-major with
 * leading dimension w_I (the concatenation of its cells (K,I) in row order).
 * Steps follow the reference line by line:
 *   factorize   blockCell_t::factorize  (symPACKMatrix2D.hpp:1052-1102) ->
 *               lapack::Potrf('U')      (src/lapack.cpp:293-306)
 *   trsm        blockCell_t::trsm       (:1104-1178), Trsm('L','U','T','N')
 *   update      blockCell_t::update     (:1180-1466): for every pair of
 *               row-supernodes J <= K of I, target cell (K,J) -= L_KI L_JI^T,
 *               scattered through the row index map (:1393-1466)
 *   LDL         blockCellLDL_t::{factorize,trsm,update} (:2096-2389) with
 *               lapack::Potf2_LDL (include/sympack/lapack.hpp:206-396):
 *               static pivots, unit L, D on the diagonal
 *   solve       forward_update / forward_update_contrib (:1486-1643),
 *               scale_contrib (:2392-2408), back_update /
 *               back_update_contrib / _tri_solve (:1645-1797),
 *               GetSolution (:6713-6757).
 * Supernodes are processed in increasing order, a valid topological order of
 * the reference's task graph (the etree is postordered).
 * Parity pins: dense LAPACK (scipy) Cholesky / closed-form LDL^T of P A P^T on
 * small matrices, and x_true of b = A x_true (tests/test_oracle_numeric.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int n, nsuper, ldl;
  const int32_t* xsuper;  /* 1-based, nsuper+1 */
  const int64_t* xlindx;  /* 1-based, nsuper+1 */
  const int32_t* lindx;   /* 1-based rows */
  int64_t* off;           /* panel offsets */
  int32_t* sup;           /* column -> supernode (0-based) */
} plan_t;

static void build(plan_t* p, int n, int nsuper, const int32_t* xsuper, const int64_t* xlindx,
                  const int32_t* lindx, int ldl) {
  p->n = n;
  p->nsuper = nsuper;
  p->ldl = ldl;
  p->xsuper = xsuper;
  p->xlindx = xlindx;
  p->lindx = lindx;
  p->off = (int64_t*)malloc(sizeof(int64_t) * (nsuper + 1));
  p->sup = (int32_t*)malloc(sizeof(int32_t) * n);
  int64_t o = 0;
  for (int I = 0; I < nsuper; ++I) {
    p->off[I] = o;
    int w = xsuper[I + 1] - xsuper[I];
    int r = (int)(xlindx[I + 1] - xlindx[I]);
    o += (int64_t)r * w;
    for (int c = xsuper[I] - 1; c < xsuper[I + 1] - 1; ++c) p->sup[c] = I;
  }
  p->off[nsuper] = o;
}

static void release(plan_t* p) {
  free(p->off);
  free(p->sup);
}

/* size of the factor storage (doubles) */
int64_t oracle_factor_size(int n, int nsuper, const int32_t* xsuper, const int64_t* xlindx) {
  (void)n;
  int64_t o = 0;
  for (int I = 0; I < nsuper; ++I)
    o += (int64_t)(xlindx[I + 1] - xlindx[I]) * (xsuper[I + 1] - xsuper[I]);
  return o;
}

/* position of row r (1-based) in the structure of supernode J (binary search) */
static int find_row(const plan_t* p, int J, int r) {
  const int32_t* b = p->lindx + (p->xlindx[J] - 1);
  int lo = 0, hi = (int)(p->xlindx[J + 1] - p->xlindx[J]) - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (b[mid] == r) return mid;
    if (b[mid] < r) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/*
 * Factorization.  Input: lower CSC of A (0-based, caller numbering), perm
 * (1-based, perm[new-1] = old).  Output L: panels (row-major, ld = w).
 * Returns 0, or the first failing pivot column (1-based, permuted order).
 */
int oracle_factor(int n, int nsuper, const int32_t* xsuper, const int64_t* xlindx,
                  const int32_t* lindx, const int32_t* perm, const int64_t* colptr,
                  const int32_t* rowind, const double* nzval, int ldl, double* L) {
  plan_t P;
  build(&P, n, nsuper, xsuper, xlindx, lindx, ldl);
  int info = 0;
  int32_t* invp = (int32_t*)malloc(sizeof(int32_t) * n);
  for (int i = 0; i < n; ++i) invp[perm[i] - 1] = i + 1;
  memset(L, 0, sizeof(double) * P.off[nsuper]); /* explicit zero fill (fill-in) */
  /* DistributeMatrix (symPACKMatrix2D.hpp:6337-6576): A(row,col) of the lower
   * triangle lands at (max,min) of the permuted indices */
  for (int c = 0; c < n; ++c)
    for (int64_t q = colptr[c]; q < colptr[c + 1]; ++q) {
      int pr = invp[rowind[q]], pc = invp[c];
      int trow = pr > pc ? pr : pc, tcol = pr > pc ? pc : pr;
      int J = P.sup[tcol - 1];
      int w = xsuper[J + 1] - xsuper[J];
      int pos = find_row(&P, J, trow);
      if (pos < 0) {
        info = -1;
        goto done;
      }
      L[P.off[J] + (int64_t)pos * w + (tcol - xsuper[J])] = nzval[q];
    }

  {
    int* map = (int*)malloc(sizeof(int) * (n + 1));
    double* buf = NULL;
    size_t bufcap = 0;
    for (int I = 0; I < nsuper; ++I) {
      const int w = xsuper[I + 1] - xsuper[I];
      const int r = (int)(xlindx[I + 1] - xlindx[I]);
      const int32_t* st = lindx + (xlindx[I] - 1);
      double* A = L + P.off[I];
      /* factorize the diagonal cell: row-major lower L, L L^T = A_II
       * (Potrf('U') on the column-major view U = L^T) */
      if (!ldl) {
        for (int j = 0; j < w; ++j) {
          double s = A[(int64_t)j * w + j];
          for (int k = 0; k < j; ++k) s -= A[(int64_t)j * w + k] * A[(int64_t)j * w + k];
          if (!(s > 0.0)) {
            if (!info) info = xsuper[I] + j;
            s = 1.0;
          }
          double d = sqrt(s);
          A[(int64_t)j * w + j] = d;
          for (int i = j + 1; i < w; ++i) {
            double t = A[(int64_t)i * w + j];
            for (int k = 0; k < j; ++k) t -= A[(int64_t)i * w + k] * A[(int64_t)j * w + k];
            A[(int64_t)i * w + j] = t / d;
          }
        }
        /* trsm: rows below, x L^T = b */
        for (int i = w; i < r; ++i) {
          double* x = A + (int64_t)i * w;
          for (int j = 0; j < w; ++j) {
            double t = x[j];
            for (int k = 0; k < j; ++k) t -= x[k] * A[(int64_t)j * w + k];
            x[j] = t / A[(int64_t)j * w + j];
          }
        }
      } else {
        /* Potf2_LDL: A = L D L^T, unit L, D on the diagonal; INFO = first
         * zero pivot (the reference then stops factoring the block). */
        for (int j = 0; j < w; ++j) {
          double s = A[(int64_t)j * w + j];
          for (int k = 0; k < j; ++k)
            s -= A[(int64_t)j * w + k] * A[(int64_t)j * w + k] * A[(int64_t)k * w + k];
          if (s == 0.0 || s != s) {
            if (!info) info = xsuper[I] + j;
            s = 1.0;
          }
          A[(int64_t)j * w + j] = s;
          for (int i = j + 1; i < w; ++i) {
            double t = A[(int64_t)i * w + j];
            for (int k = 0; k < j; ++k)
              t -= A[(int64_t)i * w + k] * A[(int64_t)j * w + k] * A[(int64_t)k * w + k];
            A[(int64_t)i * w + j] = t / s;
          }
        }
        /* trsm: unit lower, then scale columns by 1/D (blockCellLDL_t::trsm) */
        for (int i = w; i < r; ++i) {
          double* x = A + (int64_t)i * w;
          for (int j = 0; j < w; ++j) {
            double t = x[j];
            for (int k = 0; k < j; ++k) t -= x[k] * A[(int64_t)j * w + k];
            x[j] = t;
          }
          for (int j = 0; j < w; ++j) x[j] /= A[(int64_t)j * w + j];
        }
      }
      /* update: target (s_a, s_b) -= sum_k L[a][k] (D_k) L[b][k], a >= b >= w */
      const int m = r - w;
      if (m > 0) {
        size_t need = (size_t)m * w;
        if (need > bufcap) {
          free(buf);
          buf = (double*)malloc(sizeof(double) * need);
          bufcap = need;
        }
        /* W = L_off * D (LDL) or L_off (LL) */
        for (int a = 0; a < m; ++a)
          for (int k = 0; k < w; ++k)
            buf[(size_t)a * w + k] =
                A[(int64_t)(w + a) * w + k] * (ldl ? A[(int64_t)k * w + k] : 1.0);
        int b0 = 0;
        while (b0 < m) {
          int J = P.sup[st[w + b0] - 1];
          int b1 = b0;
          while (b1 < m && P.sup[st[w + b1] - 1] == J) ++b1;
          const int wJ = xsuper[J + 1] - xsuper[J];
          const int rJ = (int)(xlindx[J + 1] - xlindx[J]);
          const int32_t* stJ = lindx + (xlindx[J] - 1);
          for (int q = 0; q < rJ; ++q) map[stJ[q]] = q;
          double* T = L + P.off[J];
          for (int b = b0; b < b1; ++b) {
            const int col = st[w + b] - xsuper[J];
            const double* lb = A + (int64_t)(w + b) * w;
            for (int a = b; a < m; ++a) {
              const double* wa = buf + (size_t)a * w;
              double s = 0.0;
              for (int k = 0; k < w; ++k) s += wa[k] * lb[k];
              T[(int64_t)map[st[w + a]] * wJ + col] -= s;
            }
          }
          (void)rJ;
          b0 = b1;
        }
      }
    }
    free(buf);
    free(map);
  }
done:
  free(invp);
  release(&P);
  return info;
}

/*
 * Solve with the factor from oracle_factor.  b and x: column-major n x nrhs
 * in the caller's numbering.
 */
int oracle_solve(int n, int nsuper, const int32_t* xsuper, const int64_t* xlindx,
                 const int32_t* lindx, const int32_t* perm, int ldl, const double* L,
                 const double* b, int nrhs, double* x) {
  plan_t P;
  build(&P, n, nsuper, xsuper, xlindx, lindx, ldl);
  double* y = (double*)malloc(sizeof(double) * (size_t)n * nrhs); /* row-major n x nrhs */
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < nrhs; ++j) y[(size_t)i * nrhs + j] = b[(size_t)j * n + perm[i] - 1];
  /* forward */
  for (int I = 0; I < nsuper; ++I) {
    const int w = xsuper[I + 1] - xsuper[I], fc = xsuper[I] - 1;
    const int r = (int)(xlindx[I + 1] - xlindx[I]);
    const int32_t* st = lindx + (xlindx[I] - 1);
    const double* A = L + P.off[I];
    for (int j = 0; j < w; ++j)
      for (int c = 0; c < nrhs; ++c) {
        double t = y[(size_t)(fc + j) * nrhs + c];
        for (int k = 0; k < j; ++k) t -= A[(int64_t)j * w + k] * y[(size_t)(fc + k) * nrhs + c];
        y[(size_t)(fc + j) * nrhs + c] = ldl ? t : t / A[(int64_t)j * w + j];
      }
    for (int a = w; a < r; ++a) {
      const int row = st[a] - 1;
      for (int c = 0; c < nrhs; ++c) {
        double t = 0.0;
        for (int k = 0; k < w; ++k) t += A[(int64_t)a * w + k] * y[(size_t)(fc + k) * nrhs + c];
        y[(size_t)row * nrhs + c] -= t;
      }
    }
  }
  if (ldl)
    for (int I = 0; I < nsuper; ++I) {
      const int w = xsuper[I + 1] - xsuper[I], fc = xsuper[I] - 1;
      const double* A = L + P.off[I];
      for (int j = 0; j < w; ++j)
        for (int c = 0; c < nrhs; ++c) y[(size_t)(fc + j) * nrhs + c] /= A[(int64_t)j * w + j];
    }
  /* backward */
  for (int I = nsuper - 1; I >= 0; --I) {
    const int w = xsuper[I + 1] - xsuper[I], fc = xsuper[I] - 1;
    const int r = (int)(xlindx[I + 1] - xlindx[I]);
    const int32_t* st = lindx + (xlindx[I] - 1);
    const double* A = L + P.off[I];
    for (int k = 0; k < w; ++k)
      for (int c = 0; c < nrhs; ++c) {
        double t = 0.0;
        for (int a = w; a < r; ++a) t += A[(int64_t)a * w + k] * y[(size_t)(st[a] - 1) * nrhs + c];
        y[(size_t)(fc + k) * nrhs + c] -= t;
      }
    for (int j = w - 1; j >= 0; --j)
      for (int c = 0; c < nrhs; ++c) {
        double t = y[(size_t)(fc + j) * nrhs + c];
        for (int k = j + 1; k < w; ++k) t -= A[(int64_t)k * w + j] * y[(size_t)(fc + k) * nrhs + c];
        y[(size_t)(fc + j) * nrhs + c] = ldl ? t : t / A[(int64_t)j * w + j];
      }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < nrhs; ++j) x[(size_t)j * n + perm[i] - 1] = y[(size_t)i * nrhs + j];
  free(y);
  release(&P);
  return 0;
}

/*
 * BLAS-backed LL^T with the same storage, used as the CPU baseline: the
 * reference's per-supernode LAPACK/BLAS calls on the row-major cells
 * (Appendix of the call translation: column-major view U = L^T), with the
 * BLAS entry points passed in by the caller (scipy's OpenBLAS):
 *   factorize  Potrf('U', w, A, w)                    symPACKMatrix2D.hpp:1063,
 *                                                     src/lapack.cpp:293-306
 *   trsm       Trsm('L','U','T','N', w, m, 1, A, w, B, w)          :1120
 *   update     one Syrk('L','T', m, w, 1, B, w, 0, U, m) per supernode for
 *              all its target cells (the reference issues one Syrk/Gemm per
 *              target cell, :1288-1341), then the scatter through the
 *              relative row positions (:1393-1466), on the calling thread
 *              (the BLAS calls run on the BLAS library's own threads).
 * L holds the assembled matrix (oracle_assemble) on entry and the factor on
 * return.  Returns 0, the first failing pivot column (1-based), or -1.
 */
typedef void (*potrf_fn)(char*, int*, double*, int*, int*);
typedef void (*trsm_fn)(char*, char*, char*, char*, int*, int*, double*, double*, int*, double*, int*);
typedef void (*syrk_fn)(char*, char*, int*, int*, double*, double*, int*, double*, double*, int*);

int oracle_assemble(int n, int nsuper, const int32_t* xsuper, const int64_t* xlindx,
                    const int32_t* lindx, const int32_t* perm, const int64_t* colptr,
                    const int32_t* rowind, const double* nzval, double* L) {
  plan_t P;
  build(&P, n, nsuper, xsuper, xlindx, lindx, 0);
  int rc = 0;
  int32_t* invp = (int32_t*)malloc(sizeof(int32_t) * n);
  for (int i = 0; i < n; ++i) invp[perm[i] - 1] = i + 1;
  memset(L, 0, sizeof(double) * P.off[nsuper]);
  for (int c = 0; c < n && rc == 0; ++c)
    for (int64_t q = colptr[c]; q < colptr[c + 1]; ++q) {
      int pr = invp[rowind[q]], pc = invp[c];
      int trow = pr > pc ? pr : pc, tcol = pr > pc ? pc : pr;
      int J = P.sup[tcol - 1];
      int w = xsuper[J + 1] - xsuper[J];
      int pos = find_row(&P, J, trow);
      if (pos < 0) {
        rc = -1;
        break;
      }
      L[P.off[J] + (int64_t)pos * w + (tcol - xsuper[J])] = nzval[q];
    }
  free(invp);
  release(&P);
  return rc;
}

int oracle_factor_blas(int n, int nsuper, const int32_t* xsuper, const int64_t* xlindx,
                       const int32_t* lindx, double* L, void* potrf_p, void* trsm_p, void* syrk_p) {
  potrf_fn potrf = (potrf_fn)potrf_p;
  trsm_fn trsm = (trsm_fn)trsm_p;
  syrk_fn syrk = (syrk_fn)syrk_p;
  plan_t P;
  build(&P, n, nsuper, xsuper, xlindx, lindx, 0);
  int info = 0;
  int* map = (int*)malloc(sizeof(int) * (n + 1));
  double* U = NULL;
  size_t ucap = 0;
  char cU = 'U', cL = 'L', cT = 'T', cN = 'N';
  double one = 1.0, zero = 0.0;
  for (int I = 0; I < nsuper && info == 0; ++I) {
    int w = xsuper[I + 1] - xsuper[I];
    const int r = (int)(xlindx[I + 1] - xlindx[I]);
    int m = r - w;
    const int32_t* st = lindx + (xlindx[I] - 1);
    double* A = L + P.off[I];
    int pinfo = 0;
    potrf(&cU, &w, A, &w, &pinfo);
    if (pinfo > 0) {
      info = xsuper[I] + pinfo - 1;
      break;
    }
    if (m <= 0) continue;
    double* B = A + (int64_t)w * w;
    trsm(&cL, &cU, &cT, &cN, &w, &m, &one, A, &w, B, &w);
    if ((size_t)m * m > ucap) {
      free(U);
      ucap = (size_t)m * m;
      U = (double*)malloc(sizeof(double) * ucap);
    }
    syrk(&cL, &cT, &m, &w, &one, B, &w, &zero, U, &m);
    int b0 = 0;
    while (b0 < m) {
      const int J = P.sup[st[w + b0] - 1];
      int b1 = b0;
      while (b1 < m && P.sup[st[w + b1] - 1] == J) ++b1;
      const int wJ = xsuper[J + 1] - xsuper[J];
      const int rJ = (int)(xlindx[J + 1] - xlindx[J]);
      const int32_t* stJ = lindx + (xlindx[J] - 1);
      for (int q = 0; q < rJ; ++q) map[stJ[q]] = q;
      double* T = L + P.off[J];
      for (int b = b0; b < b1; ++b) {
        const int col = st[w + b] - xsuper[J];
        const double* ub = U + (size_t)b * m;
        for (int a = b; a < m; ++a) T[(int64_t)map[st[w + a]] * wJ + col] -= ub[a];
      }
      b0 = b1;
    }
  }
  free(U);
  free(map);
  release(&P);
  return info;
}
