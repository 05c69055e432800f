// gfx950 kernels for the supernodal LL^T / static LDL^T factorization and
// the supernodal triangular solve.
//
// Factor layout: every supernode I owns a panel of r_I x w_I entries
// (structure rows x columns), column-major with leading dimension r_I, inside
// one HBM arena.  Rows 0..w_I-1 of the panel are the diagonal block.  In LDL
// mode the diagonal of that block holds D and the strictly lower part the
// unit-lower factor, as in symPACK's blockCellLDL_t
// (include/sympack/symPACKMatrix2D.hpp:2096-2139).
//
// Kernels (one workgroup = 256 threads = 4 wave64):
//  k_potrf       diagonal-block Cholesky/LDL^T in LDS (blockCell_t::factorize,
//                symPACKMatrix2D.hpp:1052-1102).
//  k_trsm        panel TRSM of one 64-row tile below that block
//                (blockCell_t::trsm, :1104-1178).
//  k_update<0>   trailing update of a wide panel inside itself (k = 64).
//  k_update<1>   Schur update of ancestors: -P P^T (LDL: -P D P^T) on FP64 /
//                FP32 MFMA 16x16x4, scattered straight from registers into the
//                target panels through precomputed relative indices with
//                hardware atomics (blockCell_t::update, :1180-1466).
//  k_sdiag_fwd / k_soff_fwd / k_scale_d / k_soff_bwd / k_sdiag_bwd
//                supernodal forward/backward solve (forward_update[_contrib],
//                back_update[_contrib], :1486-1797).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace spk {
namespace dev {

constexpr int NB = kNB;
constexpr int TILE = kTile;
constexpr int NT = 256;

template <typename T>
struct Mfma;
// Accumulator layout (lane l, slot v) -> row i of the 16x16 block; the column
// is l%16 for both.  Measured on gfx950: the f64 instruction interleaves rows
// (i = 4v + l/16), the f32 one blocks them (i = 4(l/16) + v).
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

// Marks a register as used at this point.  Placed after a batch of
// unconditional loads it keeps the compiler from sinking each load into its
// own predicated region (which would wait for every load separately): all
// loads are issued back to back and waited for once.
template <typename T>
__device__ __forceinline__ void keep(T& v) {
  if constexpr (sizeof(T) == 8)
    asm volatile("" : "+v"(v));
  else
    asm volatile("" : "+v"(v));
}
// Compiler-only memory barrier: stops loads of later steps being hoisted
// above this point (which would raise register pressure and spill).
__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }

// ---------------------------------------------------------------------------
// Zero fill / assembly
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill_zero(T* __restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = T(0);
}

// arena[amap[e]] = scale * vals[e]; amap[e] < 0 marks a superseded duplicate.
// With equilibration, scale = s_row * s_col per entry (esc != nullptr).
template <typename T>
__global__ void k_assemble(T* __restrict__ arena, const int64_t* __restrict__ amap,
                           const double* __restrict__ vals, const double* __restrict__ esc,
                           int64_t nnz) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t dst = amap[e];
    if (dst < 0) continue;
    double v = vals[e];
    if (esc) v *= esc[e];
    arena[dst] = (T)v;
  }
}

// ---------------------------------------------------------------------------
// Diagonal block factorization (bs <= 64), one wave per block.  Lane i keeps
// row i of the block in registers; at step j the unscaled pivot column is
// broadcast through LDS and every lane applies the rank-1 update
// a_ik -= a_ij a_kj / a_jj to its own row (the same formula serves LL^T and
// LDL^T); the column is scaled at the end of the step.  Task {sn, step, -, -};
// info receives the smallest failing pivot column (1-based, permuted order).
// ---------------------------------------------------------------------------
template <typename T, bool LDL>
__global__ __launch_bounds__(64) void k_potrf(const SnInfo* __restrict__ sn,
                                               const Task4* __restrict__ tasks,
                                               T* __restrict__ arena, int* __restrict__ info) {
  __shared__ __attribute__((aligned(16))) T col[NB];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int lane = threadIdx.x;
  const int c0 = t.b * NB;
  const int bs = min(NB, s.w - c0);
  T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  // clamped, unconditional loads keep all 64 loads in flight; entries
  // outside the block's lower triangle are zero, so the update loops below
  // need no per-element guards.
  T a[NB];
  const int lr = min(lane, bs - 1);
#pragma unroll
  for (int k = 0; k < NB; ++k) a[k] = P[(int64_t)(c0 + min(k, bs - 1)) * ld + c0 + lr];
#pragma unroll
  for (int k = 0; k < NB; ++k) keep(a[k]);
#pragma unroll
  for (int k = 0; k < NB; ++k) a[k] = (lane < bs && k <= lane) ? a[k] : T(0);
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < bs) {
      col[lane] = a[j];
      __syncthreads();
      T ajj = col[j];
      const bool bad = LDL ? (ajj == T(0) || ajj != ajj) : !(ajj > T(0));
      if (bad) {
        if (lane == 0) atomicMin(info, s.fc + c0 + j + 1);
        ajj = T(1);
      }
      const T inv = T(1) / ajj;
      const T t2 = a[j] * inv;
#pragma unroll
      for (int k = j + 1; k < NB; ++k) a[k] -= t2 * col[k];
      if (LDL)
        a[j] = (lane == j) ? ajj : t2;
      else
        a[j] = (lane == j) ? sqrt(ajj) : a[j] * (T(1) / sqrt(ajj));
      __syncthreads();
    }
  }
  if (lane < bs) {
    T* prow = P + (int64_t)c0 * ld + c0 + lane;
    asm volatile("" : "+v"(prow));
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (k <= lane) prow[(int64_t)k * ld] = a[k];
  }
}

// ---------------------------------------------------------------------------
// Panel TRSM of a 256-row tile below a factored diagonal block:
// B <- B L^{-T} (LL) or B <- B L^{-T} D^{-1} (LDL, unit L).  The block is
// staged in LDS by columns; each thread then solves its own row from
// registers with broadcast LDS reads, without further barriers.
// Task {sn, step, row_tile, -}.
// ---------------------------------------------------------------------------
constexpr int TRSM_ROWS = 256;

// Body of the TRSM for a diagonal block of width bs <= W (W = 16, 32, 64):
// the unrolled loops only run over W columns.
template <typename T, bool LDL, int W>
__device__ __forceinline__ void trsm_rows(T* __restrict__ P, int64_t ld, int c0, int bs, int R0,
                                          int nrows_panel, T (*Lc)[NB], const T* invd) {
  const int tid = threadIdx.x;
  // rows past the panel work on a clamped (valid) row and do not store
  const bool active = R0 + tid < nrows_panel;
  const int row = min(R0 + tid, nrows_panel - 1);
  // columns >= bs read column bs-1 (a valid address) and are zeroed; Lc is
  // zero on and above its diagonal and invd is zero beyond bs, so the
  // substitution needs no per-element guards.
  T x[W];
  {
    const T* prow = P + (int64_t)c0 * ld + row;
#pragma unroll
    for (int j = 0; j < W; ++j) x[j] = prow[(int64_t)min(j, bs - 1) * ld];
  }
#pragma unroll
  for (int j = 0; j < W; ++j) keep(x[j]);
#pragma unroll
  for (int j = 0; j < W; ++j) x[j] = (j < bs) ? x[j] : T(0);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const T xj = LDL ? x[j] : x[j] * invd[j];
#pragma unroll
    for (int k = j + 1; k < W; ++k) x[k] -= xj * Lc[j][k];
  }
  if (!active) return;
  // a fresh row pointer: the load addresses must not stay live until here
  T* prow = P + (int64_t)c0 * ld + row;
  asm volatile("" : "+v"(prow));
  // unconditional stores in descending column order: columns >= bs write
  // zeros to column bs-1, which its own (correct) value then overwrites; a
  // guard per column would split the code into many blocks and spill
#pragma unroll
  for (int j = W - 1; j >= 0; --j) prow[(int64_t)min(j, bs - 1) * ld] = x[j] * invd[j];
}

template <typename T, bool LDL>
__global__ __launch_bounds__(NT) void k_trsm(const SnInfo* __restrict__ sn,
                                              const Task4* __restrict__ tasks,
                                              T* __restrict__ arena) {
  __shared__ __attribute__((aligned(16))) T Lc[NB][NB];  // Lc[j][k] = L(k, j), k > j
  __shared__ T invd[NB];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int tid = threadIdx.x;
  const int c0 = t.b * NB;
  const int bs = min(NB, s.w - c0);
  T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int R0 = c0 + bs + t.c * TRSM_ROWS;
  for (int idx = tid; idx < NB * NB; idx += NT) {
    const int k = idx % NB, j = idx / NB;  // column j, row k: coalesced down the column
    const T v = P[(int64_t)(c0 + min(j, bs - 1)) * ld + c0 + min(k, bs - 1)];
    Lc[j][k] = (k < bs && j < bs && k > j) ? v : T(0);
    if (k == j) invd[j] = (j < bs) ? T(1) / v : T(0);
  }
  __syncthreads();
  if (bs <= 16)
    trsm_rows<T, LDL, 16>(P, ld, c0, bs, R0, s.r, Lc, invd);
  else if (bs <= 32)
    trsm_rows<T, LDL, 32>(P, ld, c0, bs, R0, s.r, Lc, invd);
  else
    trsm_rows<T, LDL, 64>(P, ld, c0, bs, R0, s.r, Lc, invd);
}

// ---------------------------------------------------------------------------
// MFMA tile update.  The 64x64 output tile C[a][b] = sum_k Pa[a][k] Pb[b][k]
// (times D[k] for LDL) is computed by 4 waves, each owning a 32x32 quarter as
// 2x2 blocks of 16x16x4 MFMA.  MFMA operand A is the b side and operand B the
// a side, so an accumulator element (lane l, slot v) holds b = Mfma<T>::row(l,v)
// and a = l%16 within its block: 16 consecutive lanes own 16 consecutive
// target rows, which keeps the scatter coalesced.
// MODE 0: in-panel trailing update of step t.b (direct read-modify-write).
// MODE 1: Schur update of ancestor panels (atomic scatter).
// ---------------------------------------------------------------------------
template <typename T, int MODE, bool LDL>
__global__ __launch_bounds__(NT) void k_update(const SnInfo* __restrict__ sn,
                                                const Task4* __restrict__ tasks,
                                                T* __restrict__ arena,
                                                const int64_t* __restrict__ tcol,
                                                const int64_t* __restrict__ tmap,
                                                const int32_t* __restrict__ rel) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int KC = 32;
  constexpr int LDS_LD = TILE + 16;
  __shared__ T As[KC][LDS_LD];
  __shared__ T Bs[KC][LDS_LD];
  __shared__ T Dk[KC];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  int kbeg, K, arow0, brow0, alim, blim, ti, tj;
  if (MODE == 0) {
    kbeg = t.b * NB;
    K = min(NB, s.w - kbeg);
    const int rb = kbeg + K;
    ti = t.c;
    tj = t.d;
    arow0 = rb + ti * TILE;
    brow0 = rb + tj * TILE;
    alim = s.r;
    blim = s.w;
  } else {
    kbeg = 0;
    K = s.w;
    ti = t.b;
    tj = t.c;
    arow0 = s.w + ti * TILE;
    brow0 = s.w + tj * TILE;
    alim = s.r;
    blim = s.r;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wb = wave >> 1, wa = wave & 1;
  v4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};

  // loader mapping: a wave reads 64 consecutive rows of one k column
  const int lrow = tid & 63, lk = tid >> 6;
  const bool arow_ok = arow0 + lrow < alim;
  const bool brow_ok = brow0 + lrow < blim;
  const int arow = min(arow0 + lrow, alim - 1), brow = min(brow0 + lrow, blim - 1);
  constexpr int NQ = KC / 4;
  T ra[NQ], rb[NQ];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kg = kc + lk + 4 * q;
      const T* col = P + (int64_t)(kbeg + min(kg, K - 1)) * ld;
      // clamped addresses hold finite values: masking by multiplication
      // keeps the loads unconditional (a select would be turned into a
      // predicated load with its own wait)
      ra[q] = col[arow] * T((kg < K) & arow_ok);
      rb[q] = col[brow] * T((kg < K) & brow_ok);
    }
  };
  fetch(0);
  for (int kc = 0; kc < K; kc += KC) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      As[lk + 4 * q][lrow] = ra[q];
      Bs[lk + 4 * q][lrow] = rb[q];
    }
    if (LDL && tid < KC) {
      const int kg = kc + tid;
      Dk[tid] = (kg < K) ? P[(int64_t)(kbeg + kg) * ld + kbeg + kg] : T(0);
    }
    __syncthreads();
    if (kc + KC < K) fetch(kc + KC);  // next chunk in flight during the MFMAs
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      T bf[2], af[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
        bf[bi] = Bs[k][wb * 32 + bi * 16 + (lane & 15)];
        if (LDL) bf[bi] *= Dk[k];
      }
#pragma unroll
      for (int aj = 0; aj < 2; ++aj) af[aj] = As[k][wa * 32 + aj * 16 + (lane & 15)];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) acc[bi][aj] = Mfma<T>::mma(bf[bi], af[aj], acc[bi][aj]);
    }
    __syncthreads();
  }

  // Epilogue: every load first (clamped, unconditional, one wait), then the
  // predicated stores / atomics.
  if (MODE == 0) {
    T cur[2][2][4];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int aj = 0; aj < 2; ++aj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int bg = brow0 + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
          const int ag = arow0 + wa * 32 + aj * 16 + (lane & 15);
          cur[bi][aj][v] = P[(int64_t)min(bg, s.w - 1) * ld + min(ag, s.r - 1)];
        }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int aj = 0; aj < 2; ++aj)
#pragma unroll
        for (int v = 0; v < 4; ++v) keep(cur[bi][aj][v]);
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int aj = 0; aj < 2; ++aj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int bg = brow0 + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
          const int ag = arow0 + wa * 32 + aj * 16 + (lane & 15);
          if (ag < s.r && bg < s.w && ag >= bg)
            P[(int64_t)bg * ld + ag] = cur[bi][aj][v] - acc[bi][aj][v];
        }
  } else {
    const int m = s.r - s.w;
    int64_t tc[2][4], tm[2][4];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int boff = tj * TILE + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
        const int bc = min(boff, m - 1);
        tc[bi][v] = tcol[s.upd + bc];
        tm[bi][v] = tmap[s.upd + bc];
      }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        keep(tc[bi][v]);
        keep(tm[bi][v]);
      }
    int32_t rl[2][4][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int boff = tj * TILE + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
        const int bc = min(boff, m - 1);
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) {
          const int aoff = ti * TILE + wa * 32 + aj * 16 + (lane & 15);
          // rel is defined for rows a >= the first row of b's target group
          const int ac = min(max(aoff, bc), m - 1);
          rl[bi][v][aj] = rel[tm[bi][v] + ac];
        }
      }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) keep(rl[bi][v][aj]);
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int boff = tj * TILE + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) {
          const int aoff = ti * TILE + wa * 32 + aj * 16 + (lane & 15);
          if (boff < m && aoff < m && aoff >= boff)
            atomic_add(arena + tc[bi][v] + rl[bi][v][aj], -acc[bi][aj][v]);
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Solve.  Vectors are row-major n x nrhs in the permuted numbering.
// Forward:  X_I = L_II^{-1} Y_I  (k_sdiag_fwd), then Y[struct rows] -= L_off X_I
//           (k_soff_fwd, atomics).  LDL then scales by D^{-1} (k_scale_d).
// Backward: X_I -= L_off^T X[struct rows] (k_soff_bwd, atomics into X_I),
//           then X_I = L_II^{-T} X_I (k_sdiag_bwd).
// ---------------------------------------------------------------------------
constexpr int JC = 64;  // right-hand sides per LDS chunk

__device__ __forceinline__ double readlane_t(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane_t(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Triangular solve of one diagonal block (bs <= 64) for right-hand sides
// handled one per wave: lane i holds x_i, the pivot value is broadcast with
// readlane and every lane applies one fused multiply-add per step.
// Forward: L x = z; backward: L^T x = z.  Lk[j][k] = L(c0+k, c0+j) for k > j
// (zero elsewhere); invd = 1/L_jj (LL) or 1 (LDL, unit diagonal).
template <typename T, bool BWD>
__device__ __forceinline__ void block_solve_waves(T* __restrict__ X, int64_t rowbase, int nrhs,
                                                  int bs, T (*Lk)[NB + 1], const T* invd) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = min(lane, bs - 1);
  for (int jj = wave; jj < nrhs; jj += NT / 64) {
    T* col = X + rowbase * nrhs + jj;
    T xi = col[(int64_t)lr * nrhs];
    xi = (lane < bs) ? xi : T(0);
    if (!BWD) {
      for (int j = 0; j < bs; ++j) {
        const T xj = readlane_t(xi, j) * invd[j];
        xi = (lane == j) ? xj : xi - Lk[j][lane] * xj;
      }
    } else {
      for (int j = bs - 1; j >= 0; --j) {
        const T xj = readlane_t(xi, j) * invd[j];
        xi = (lane == j) ? xj : xi - Lk[lane][j] * xj;
      }
    }
    if (lane < bs) col[(int64_t)lane * nrhs] = xi;
  }
}

// Stage diagonal block c0 of a panel in LDS: Lk[j][k] = L(c0+k, c0+j) for
// k > j (zero elsewhere) and invd[j] = 1/L(c0+j,c0+j) (LL) or 1 (LDL).
template <typename T, bool LDL>
__device__ __forceinline__ void stage_diag_block(const T* __restrict__ P, int64_t ld, int c0, int bs,
                                                 T (*Lk)[NB + 1], T* invd) {
  for (int idx = threadIdx.x; idx < NB * NB; idx += NT) {
    const int k = idx % NB, j = idx / NB;
    const T v = P[(int64_t)(c0 + min(j, bs - 1)) * ld + c0 + min(k, bs - 1)];
    Lk[j][k] = (k < bs && j < bs && k > j) ? v : T(0);
    if (k == j) invd[j] = (j < bs) ? (LDL ? T(1) : T(1) / v) : T(0);
  }
}

// Forward diagonal solve of a supernode: X_I = L_II^{-1} Y_I, blocked by 64
// columns; task {sn}.  Threads own right-hand sides for the block solves and
// (row, rhs) pairs for the updates of the supernode's remaining rows.
template <typename T, bool LDL>
__global__ __launch_bounds__(NT) void k_sdiag_fwd(const SnInfo* __restrict__ sn,
                                                   const Task4* __restrict__ tasks,
                                                   const T* __restrict__ arena,
                                                   const T* __restrict__ Y, T* __restrict__ X,
                                                   int nrhs) {
  __shared__ T Lk[NB][NB + 1];
  __shared__ T invd[NB];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int tid = threadIdx.x;
  const T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int64_t base = (int64_t)s.fc * nrhs;
  for (int64_t i = tid; i < (int64_t)s.w * nrhs; i += NT) X[base + i] = Y[base + i];
  __syncthreads();
  for (int c0 = 0; c0 < s.w; c0 += NB) {
    const int bs = min(NB, s.w - c0);
    stage_diag_block<T, LDL>(P, ld, c0, bs, Lk, invd);
    __syncthreads();
    block_solve_waves<T, false>(X, (int64_t)s.fc + c0, nrhs, bs, Lk, invd);
    __syncthreads();
    // rows of the supernode below this block: X_i -= L(i, block) x_block;
    // fixed-length unrolled dot products (clamped columns, zero weights
    // beyond bs) keep all loads of a row in flight.
    const int below = s.w - c0 - bs;
    for (int idx = tid; idx < below * nrhs; idx += NT) {
      const int i = c0 + bs + idx / nrhs, jj = idx % nrhs;
      const T* lrow = P + (int64_t)c0 * ld + i;
      const T* xb = X + base + (int64_t)c0 * nrhs + jj;
      T lv[NB], xv[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        lv[k] = lrow[(int64_t)min(k, bs - 1) * ld];
        xv[k] = xb[(int64_t)min(k, bs - 1) * nrhs];
      }
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < NB; ++k) acc += (k < bs) ? lv[k] * xv[k] : T(0);
      X[base + (int64_t)i * nrhs + jj] -= acc;
    }
    __syncthreads();
  }
}

// Off-diagonal tile (rows rt*64.., columns kt*64.. of the panel below the
// diagonal block).  Task {sn, row_tile, k_tile, -}.
template <typename T>
__global__ __launch_bounds__(NT) void k_soff_fwd(const SnInfo* __restrict__ sn,
                                                  const Task4* __restrict__ tasks,
                                                  const T* __restrict__ arena,
                                                  const int32_t* __restrict__ lindx0,
                                                  const T* __restrict__ X, T* __restrict__ Y,
                                                  int nrhs) {
  __shared__ T Ps[NB][TILE + 1];  // Ps[k][a]
  __shared__ T Xs[NB][JC + 1];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int tid = threadIdx.x;
  const T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int m = s.r - s.w;
  const int a0 = t.b * TILE, k0 = t.c * NB;
  const int na = min(TILE, m - a0), nk = min(NB, s.w - k0);
  for (int idx = tid; idx < NB * TILE; idx += NT) {
    int a = idx % TILE, k = idx / TILE;
    Ps[k][a] = (a < na && k < nk) ? P[(int64_t)(k0 + k) * ld + s.w + a0 + a] : T(0);
  }
  const int a = tid % TILE, g = tid / TILE;
  const int row = (a < na) ? lindx0[s.lx + s.w + a0 + a] : 0;
  for (int j0 = 0; j0 < nrhs; j0 += JC) {
    const int jc = min(JC, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < nk * jc; idx += NT) {
      int k = idx / jc, jj = idx % jc;
      Xs[k][jj] = X[(int64_t)(s.fc + k0 + k) * nrhs + j0 + jj];
    }
    __syncthreads();
    if (a < na) {
      for (int jj = g; jj < jc; jj += 4) {
        T acc = T(0);
        for (int k = 0; k < nk; ++k) acc += Ps[k][a] * Xs[k][jj];
        atomic_add(Y + (int64_t)row * nrhs + j0 + jj, -acc);
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_soff_bwd(const SnInfo* __restrict__ sn,
                                                  const Task4* __restrict__ tasks,
                                                  const T* __restrict__ arena,
                                                  const int32_t* __restrict__ lindx0,
                                                  T* __restrict__ X, int nrhs) {
  __shared__ T Ps[NB][TILE + 1];  // Ps[k][a]
  __shared__ T Xs[TILE][JC + 1];  // gathered solution rows
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int tid = threadIdx.x;
  const T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int m = s.r - s.w;
  const int a0 = t.b * TILE, k0 = t.c * NB;
  const int na = min(TILE, m - a0), nk = min(NB, s.w - k0);
  for (int idx = tid; idx < NB * TILE; idx += NT) {
    int a = idx % TILE, k = idx / TILE;
    Ps[k][a] = (a < na && k < nk) ? P[(int64_t)(k0 + k) * ld + s.w + a0 + a] : T(0);
  }
  const int k = tid % NB, g = tid / NB;
  for (int j0 = 0; j0 < nrhs; j0 += JC) {
    const int jc = min(JC, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < na * jc; idx += NT) {
      int a = idx / jc, jj = idx % jc;
      int row = lindx0[s.lx + s.w + a0 + a];
      Xs[a][jj] = X[(int64_t)row * nrhs + j0 + jj];
    }
    __syncthreads();
    if (k < nk) {
      for (int jj = g; jj < jc; jj += 4) {
        T acc = T(0);
        for (int a = 0; a < na; ++a) acc += Ps[k][a] * Xs[a][jj];
        atomic_add(X + (int64_t)(s.fc + k0 + k) * nrhs + j0 + jj, -acc);
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_scale_d(const SnInfo* __restrict__ sn,
                                                 const Task4* __restrict__ tasks,
                                                 const T* __restrict__ arena, T* __restrict__ X,
                                                 int nrhs) {
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const T* __restrict__ P = arena + s.off;
  for (int64_t idx = threadIdx.x; idx < (int64_t)s.w * nrhs; idx += NT) {
    int i = (int)(idx / nrhs);
    X[(int64_t)s.fc * nrhs + idx] /= P[(int64_t)i * s.r + i];
  }
}

// Backward diagonal solve of a supernode: X_I = L_II^{-T} X_I, blocks from
// the last one; task {sn}.
template <typename T, bool LDL>
__global__ __launch_bounds__(NT) void k_sdiag_bwd(const SnInfo* __restrict__ sn,
                                                   const Task4* __restrict__ tasks,
                                                   const T* __restrict__ arena,
                                                   T* __restrict__ X, int nrhs) {
  __shared__ T Lk[NB][NB + 1];
  __shared__ T invd[NB];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int tid = threadIdx.x;
  const T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int64_t base = (int64_t)s.fc * nrhs;
  const int nb = (s.w + NB - 1) / NB;
  for (int blk = nb - 1; blk >= 0; --blk) {
    const int c0 = blk * NB;
    const int bs = min(NB, s.w - c0);
    // X_block -= L(rows below the block inside the supernode, block)^T x_below
    const int below = s.w - c0 - bs;
    for (int idx = tid; idx < bs * nrhs; idx += NT) {
      const int k = idx / nrhs, jj = idx % nrhs;
      const T* lcol = P + (int64_t)(c0 + k) * ld + c0 + bs;
      const T* xb = X + base + (int64_t)(c0 + bs) * nrhs + jj;
      T acc = T(0);
      for (int i0 = 0; i0 < below; i0 += NB) {
        const int cnt = min(NB, below - i0);
        T lv[NB], xv[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          lv[i] = lcol[i0 + min(i, cnt - 1)];
          xv[i] = xb[(int64_t)(i0 + min(i, cnt - 1)) * nrhs];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) acc += (i < cnt) ? lv[i] * xv[i] : T(0);
      }
      X[base + (int64_t)(c0 + k) * nrhs + jj] -= acc;
    }
    stage_diag_block<T, LDL>(P, ld, c0, bs, Lk, invd);
    __syncthreads();
    block_solve_waves<T, true>(X, (int64_t)s.fc + c0, nrhs, bs, Lk, invd);
    __syncthreads();
  }
}

// b (column-major, original numbering) -> Y (row-major, permuted numbering)
template <typename T>
__global__ void k_gather_rhs(const double* __restrict__ b, T* __restrict__ Y,
                             const int32_t* __restrict__ perm0, int n, int nrhs,
                             const double* __restrict__ esc_row) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(idx / nrhs), j = (int)(idx % nrhs);
    int o = perm0[i];
    double v = b[(int64_t)j * n + o];
    if (esc_row) v *= esc_row[o];
    Y[idx] = (T)v;
  }
}

// X (row-major, permuted) -> x (column-major, original numbering)
template <typename T>
__global__ void k_scatter_sol(const T* __restrict__ X, double* __restrict__ x,
                              const int32_t* __restrict__ perm0, int n, int nrhs,
                              const double* __restrict__ esc_row, int accumulate) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(idx / nrhs), j = (int)(idx % nrhs);
    int o = perm0[i];
    double v = (double)X[idx];
    if (esc_row) v *= esc_row[o];
    double* dst = x + (int64_t)j * n + o;
    if (accumulate)
      *dst += v;
    else
      *dst = v;
  }
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - A x with A given by its lower triangle (CSC, 0-based)
// in the original numbering, as check_solution does (driver/utils.hpp:95-142):
// one workgroup-stride loop per column, symmetric contributions by atomics.
// ---------------------------------------------------------------------------
__global__ void k_residual(const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
                           const double* __restrict__ val, const double* __restrict__ x,
                           double* __restrict__ r, int n, int nrhs) {
  int c = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (c >= n) return;
  for (int j = 0; j < nrhs; ++j) {
    const double* xj = x + (int64_t)j * n;
    double* rj = r + (int64_t)j * n;
    double xc = xj[c];
    double acc = 0.0;
    for (int64_t p = colptr[c] + lane; p < colptr[c + 1]; p += 64) {
      int row = rowind[p];
      double v = val[p];
      if (row != c) {
        unsafeAtomicAdd(rj + row, -v * xc);  // A(row,c) x(c)
        acc += v * xj[row];             // A(c,row) x(row)
      } else {
        acc += v * xc;
      }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane == 0) unsafeAtomicAdd(rj + c, -acc);
  }
}

}  // namespace dev
}  // namespace spk
