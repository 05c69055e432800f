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
//  k_diag        Cholesky/LDL^T of a diagonal block of up to 160 columns in
//                LDS (blockCell_t::factorize, symPACKMatrix2D.hpp:1052-1102).
//  k_ptrsm       panel TRSM of 64 rows below that block on MFMA with the
//                inverted 16x16 diagonal tiles (blockCell_t::trsm, :1104-1178).
//  k_diagtrsm    both, one workgroup per 64-row tile (each factors the block).
//  k_update<0>   trailing update of a panel wider than 160 inside itself.
//  k_update<1>   Schur update of ancestors: -P P^T (LDL: -P D P^T) on FP64 /
//                FP32 MFMA 16x16x4, scattered straight from registers into the
//                target panels through precomputed relative indices with
//                hardware atomics (blockCell_t::update, :1180-1466).
//  k_sdiag4 / k_sdiag3 / k_sdiag_fwd / k_soffm_fwd / k_scale_d / k_soffm_bwd /
//  k_sdiag_bwd   supernodal forward/backward solve (forward_update[_contrib],
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

template <typename T>
__global__ void k_copy(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
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

// Relative indices of the Schur scatter (the reference computes them per
// update task, symPACKMatrix2D.hpp:1403-1439): for every group of rows of
// supernode I that are columns of ancestor J, the position inside struct(J)
// of each of I's off-diagonal rows from the group's first one on, found by
// binary search in J's sorted structure.  One workgroup per group (grid
// stride); *bad is set when a row is missing from struct(J), which a
// symbolic structure closed under elimination never allows.
__global__ __launch_bounds__(256) void k_build_rel(const SnInfo* __restrict__ sn,
                                                   const int32_t* __restrict__ lindx0,
                                                   const RelGroup* __restrict__ groups,
                                                   int64_t ngroups, int32_t* __restrict__ rel,
                                                   int64_t* __restrict__ tcol,
                                                   int64_t* __restrict__ tmap,
                                                   int* __restrict__ bad) {
  for (int64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    const RelGroup g = groups[gi];
    const SnInfo si = sn[g.I], sj = sn[g.J];
    const int m = si.r - si.w;
    const int32_t* __restrict__ rows = lindx0 + si.lx + si.w;
    const int32_t* __restrict__ sJ = lindx0 + sj.lx;
    // scatter bases of the group's rows b (columns of J): the arena offset
    // of target column b and the rel index of row a = 0
    for (int b = g.gb + (int)threadIdx.x; b < g.ge; b += (int)blockDim.x) {
      tcol[si.upd + b] = sj.off + (int64_t)(rows[b] - sj.fc) * sj.r;
      tmap[si.upd + b] = g.base - g.gb;
    }
    for (int a = g.gb + (int)threadIdx.x; a < m; a += (int)blockDim.x) {
      const int32_t row = rows[a];
      int lo = 0, hi = sj.r;  // first position whose row is >= row
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sJ[mid] < row)
          lo = mid + 1;
        else
          hi = mid;
      }
      const bool found = lo < sj.r && sJ[lo] == row;
      if (!found) atomicOr(bad, 1);
      rel[g.base + (a - g.gb)] = found ? lo : 0;
    }
  }
}

// Schur task lists from the per-supernode records: one workgroup per
// supernode writes its 64 x 64 tiles (plan order and longest-first
// positions) and its 128 x 64 tiles; a thread per row tile.
__global__ __launch_bounds__(256) void k_emit_schur(const SchurEmit* __restrict__ emit, int64_t nemit,
                                                    Task4* __restrict__ t0, Task4* __restrict__ t1,
                                                    Task4* __restrict__ t2, Task4* __restrict__ t3) {
  for (int64_t q = blockIdx.x; q < nemit; q += gridDim.x) {
    const SchurEmit e = emit[q];
    for (int ti = (int)threadIdx.x; ti < e.nt; ti += (int)blockDim.x) {
      const int64_t at = schur_tiles_before(ti, e.ntj);
      const int ntj = min(ti + 1, e.ntj);
      for (int tj = 0; tj < ntj; ++tj) {
        const Task4 t{e.I, ti, tj, e.d};
        t0[e.first[0] + at + tj] = t;
        t1[e.first[1] + at + tj] = t;
      }
    }
    for (int ti = (int)threadIdx.x; ti < e.ntw; ti += (int)blockDim.x) {
      const int64_t at = schur_wide_before(ti, e.ntj);
      const int ntj = min(2 * ti + 2, e.ntj);
      for (int tj = 0; tj < ntj; ++tj) {
        const Task4 t{e.I, ti, tj, e.d};
        t2[e.first[2] + at + tj] = t;
        t3[e.first[3] + at + tj] = t;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Scalar helpers for the pivots: hardware reciprocal / reciprocal square root
// refined by Newton-Raphson (y <- y(3 - a y^2)/2, y <- y(2 - a y)), two steps
// for double, one for float.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rsqrt_nr(double a) {
  double y = __builtin_amdgcn_rsq(a);
  const double h = 0.5 * a;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  return y;
}
__device__ __forceinline__ float rsqrt_nr(float a) {
  float y = __builtin_amdgcn_rsqf(a);
  y = y * (1.5f - 0.5f * a * y * y);
  return y;
}
__device__ __forceinline__ double rcp_nr(double a) {
  double y = __builtin_amdgcn_rcp(a);
  y = y * (2.0 - a * y);
  y = y * (2.0 - a * y);
  return y;
}
__device__ __forceinline__ float rcp_nr(float a) {
  float y = __builtin_amdgcn_rcpf(a);
  y = y * (2.0f - a * y);
  return y;
}
// Reciprocal from the hardware estimate y0 and the correction factor s:
// with e = 1 - a y0, 1/a = y0 (1 + e + e^2 + ...), so y0 (1 + s) with
// s = e + e^2 leaves a relative error of e^3.  The gfx950 f64 estimate has
// |e| < 2^-24 (measured over 2^20 arguments spanning 2^-30..2^30), hence
// e^3 < 2^-72: full double precision.  The f32 estimate is already within
// an ulp or two: s = e, one step.
__device__ __forceinline__ double rcp_est(double a) { return __builtin_amdgcn_rcp(a); }
__device__ __forceinline__ float rcp_est(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double rcp_corr(double e) { return fma(e, e, e); }
__device__ __forceinline__ float rcp_corr(float e) { return e; }

__device__ __forceinline__ double readlane_t(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane_t(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ---------------------------------------------------------------------------
// Diagonal blocks are held in LDS as the lower triangle of 16x16 tiles:
// tile (I,J), I >= J, at slot I(I+1)/2 + J, column-major with column stride
// 17 (conflict-free both along columns and along rows).
// ---------------------------------------------------------------------------
constexpr int SUB = kSub;
constexpr int TSTR = SUB + 1;
// Trailing tiles (I, K) = (i+1, k+1), 0 <= k <= i <= nb-2, of a diagonal
// block with nb 16-column blocks, ordered by k descending (then i):
// kOrd[kOrdOff[nb] + t] = (i << 4) | k.  The tiles still to be updated at
// block step J (k >= J) are then a prefix of the list.
__constant__ unsigned char kOrdOff[12] = {0, 0, 0, 1, 4, 10, 20, 35, 56, 84, 120, 165};
__constant__ unsigned char kOrd[165] = {
    0x00, 0x11, 0x00, 0x10, 0x22, 0x11, 0x21, 0x00, 0x10, 0x20, 0x33, 0x22, 0x32, 0x11, 0x21,
    0x31, 0x00, 0x10, 0x20, 0x30, 0x44, 0x33, 0x43, 0x22, 0x32, 0x42, 0x11, 0x21, 0x31, 0x41,
    0x00, 0x10, 0x20, 0x30, 0x40, 0x55, 0x44, 0x54, 0x33, 0x43, 0x53, 0x22, 0x32, 0x42, 0x52,
    0x11, 0x21, 0x31, 0x41, 0x51, 0x00, 0x10, 0x20, 0x30, 0x40, 0x50, 0x66, 0x55, 0x65, 0x44,
    0x54, 0x64, 0x33, 0x43, 0x53, 0x63, 0x22, 0x32, 0x42, 0x52, 0x62, 0x11, 0x21, 0x31, 0x41,
    0x51, 0x61, 0x00, 0x10, 0x20, 0x30, 0x40, 0x50, 0x60, 0x77, 0x66, 0x76, 0x55, 0x65, 0x75,
    0x44, 0x54, 0x64, 0x74, 0x33, 0x43, 0x53, 0x63, 0x73, 0x22, 0x32, 0x42, 0x52, 0x62, 0x72,
    0x11, 0x21, 0x31, 0x41, 0x51, 0x61, 0x71, 0x00, 0x10, 0x20, 0x30, 0x40, 0x50, 0x60, 0x70,
    0x88, 0x77, 0x87, 0x66, 0x76, 0x86, 0x55, 0x65, 0x75, 0x85, 0x44, 0x54, 0x64, 0x74, 0x84,
    0x33, 0x43, 0x53, 0x63, 0x73, 0x83, 0x22, 0x32, 0x42, 0x52, 0x62, 0x72, 0x82, 0x11, 0x21,
    0x31, 0x41, 0x51, 0x61, 0x71, 0x81, 0x00, 0x10, 0x20, 0x30, 0x40, 0x50, 0x60, 0x70, 0x80};
constexpr int TSZ = SUB * TSTR;
__host__ __device__ constexpr int tslot(int I, int J) { return I * (I + 1) / 2 + J; }

// ---------------------------------------------------------------------------
// Factorization of one diagonal block (bs <= 160 columns of a panel step),
// one workgroup per block, the block resident in LDS
// (blockCell_t::factorize, symPACKMatrix2D.hpp:1052-1102; LDL^T:
// blockCellLDL_t::factorize, :2096-2116, lapack.hpp:206-617).
// Right-looking over 16-column block columns J:
//  1. every wave factors block column J lane-per-row: lanes 0-15 hold the
//     16 rows of the diagonal tile (each wave keeps its own copy, so the
//     pivot columns are broadcast through a per-wave LDS buffer without
//     barriers), lanes 16-63 hold 48 rows below it;
//  2. the trailing lower tiles get A_IK -= L_IJ L_KJ^T (LDL: L_IJ D_J L_KJ^T)
//     on MFMA 16x16x4, each wave's tiles one after another with the next
//     tile's operands read from LDS ahead of the current tile's MFMAs.
// Finally the 16x16 diagonal tiles are inverted (for the panel TRSM), all in
// one pass with 16 lanes per tile.
// Rows/columns beyond bs are padded with an identity block.  Task
// {sn, step, inverse-tile offset, -}; info receives the smallest failing
// pivot column (1-based, permuted order).
// ---------------------------------------------------------------------------
// Slots 0..na-1 of a wave's trailing tiles as nested conditionals (slot q+1
// is reached only from slot q): the operands of slot q+1 are loaded before
// the MFMAs of slot q, and the compiler's LDS wait counts stay exact.
template <int Q, int NS>
struct SlotChain {
  template <class Load, class Mma>
  static __device__ __forceinline__ void run(int na, Load& load, Mma& mma) {
    if constexpr (Q < NS) {
      if (Q < na) {
        if constexpr (Q + 1 < NS) load(Q + 1);
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the MFMAs
        mma(Q);
        SlotChain<Q + 1, NS>::run(na, load, mma);
      }
    }
  }
};

// LDS of the block factor (and of the fused block factor + TRSM)
template <typename T, int NBM>
struct DiagLds {
  T tl[NBM * (NBM + 1) / 2 * TSZ];
  T dsh[SUB];
  T ysh[NBM * SUB];  // LL: 1/L_jj of every column; LDL (fused): 1/D_j
  T colb[NT / 64][2][SUB];  // per wave: the current diagonal-row column, double-buffered
  int store;  // fused: this workgroup writes the factored block back
};

// One block column J of a diagonal block held in LDS (tiles (I, J), I >= J,
// of tl), factored lane per row by one wave: lanes 0-15 hold the 16 rows of
// the diagonal tile (every wave that calls this keeps its own copy, so the
// pivot columns are broadcast through the wave's LDS buffer cb without
// barriers), lanes 16-63 hold 48 rows below it, starting 48 * wave rows
// down.  On return a[] holds the lane's factored row (LL: L_ij; LDL: l_ij
// below the diagonal, d_j on it) and the rows below are written back; the
// diagonal tile is not (every wave reads it unfactored: the caller writes it
// after a barrier through the returned row pointer).
template <typename T, bool LDL>
__device__ __forceinline__ T* chain_column(T* __restrict__ tl, T (*__restrict__ cb)[SUB], int J, int nb,
                                           int wave, int lane, T (&a)[SUB]) {
  const bool diag_lane = lane < SUB;
  const int rb = 48 * wave + lane - SUB;  // row below the diagonal tile (lanes >= 16)
  const int Ib = J + 1 + (diag_lane ? 0 : rb / SUB);
  const bool row_ok = diag_lane || (Ib < nb);
  const int I = diag_lane ? J : min(Ib, nb - 1);
  const int rr = diag_lane ? lane : rb % SUB;
  T* trow = tl + tslot(I, J) * TSZ + rr;
#pragma unroll
  for (int c = 0; c < SUB; ++c) a[c] = trow[c * TSTR];
  // every wave reads the diagonal tile: wave 0 writes the factored one
  // back only after the barrier that ends this phase
  // Column j of the diagonal rows (still unscaled) is broadcast through
  // this wave's LDS buffer: LDS operations of one wave complete in order,
  // so no barrier is needed between the write and the reads.  Column j+1
  // is final on the diagonal rows as soon as step j has updated it (the
  // first thing step j does), so it is published and read back in step j,
  // a step ahead of its use: the LDS round trip stays off the wave's path.
  T u[SUB];  // u[k], k >= j+3: column j of diagonal row k
  if (diag_lane) cb[0][lane] = a[0];
  compiler_fence();
#pragma unroll
  for (int k = 3; k < SUB; ++k) u[k] = cb[0][k];
  compiler_fence();
#pragma unroll
  for (int j = 0; j < SUB; ++j) {
    // The dependency chain from one pivot to the next runs through
    // readlane(pivot) -> 1/pivot -> the next row's diagonal entry; the
    // next row's entry comes by readlane, and the scaling 1/sqrt(pivot)
    // stays off the chain.
    const T piv = readlane_t(a[j], j);
    // u[j+1] and u[j+2] feed the next two pivots: readlane
    if (j + 1 < SUB) u[j + 1] = readlane_t(a[j], j + 1);
    if (j + 2 < SUB) u[j + 2] = readlane_t(a[j], j + 2);
    // 1/piv from the hardware estimate y0 and e = 1 - piv y0:
    // inv = y0 (1 + s) with s = e + e^2 (f64; |e| < 2^-24, so the error is
    // e^3) or s = e (f32)
    const T y0 = rcp_est(piv);
    const T e = fma(-piv, y0, T(1));
    const T s = rcp_corr(e);
    const T inv = fma(y0, s, y0);
    // a_k -= (a_j / piv) u_k.  The entry that feeds the next pivot is
    // a - t (1 + s) with t = a_j u_{j+1} y0: t and e are formed side by
    // side, so it is three dependent operations after the estimate
    if (j + 1 < SUB) {
      const T t = (a[j] * u[j + 1]) * y0;
      a[j + 1] = fma(-t, s, a[j + 1] - t);
    }
    // column j+1 of the diagonal rows is final: publish it, read the next
    // step's entries (k >= j+4)
    T un[SUB];
    if (j + 1 < SUB) {
      if (diag_lane) cb[(j + 1) & 1][lane] = a[j + 1];
      compiler_fence();
#pragma unroll
      for (int k = j + 4; k < SUB; ++k) un[k] = cb[(j + 1) & 1][k];
      compiler_fence();
    }
    const T l = a[j] * inv;
#pragma unroll
    for (int k = j + 2; k < SUB; ++k) a[k] -= l * u[k];
    // LL: L_ij = a_ij / sqrt(d_j), with 1/sqrt(d_j) formed from the
    // uniform pivot off the chain; LDL stores l = a_ij / d_j and d_j
    if (LDL) a[j] = (lane == j) ? piv : l;
    else a[j] = a[j] * rsqrt_nr(piv);
#pragma unroll
    for (int k = j + 4; k < SUB; ++k) u[k] = un[k];
  }
  // rows below: every valid lane (the trailing update reads them)
  if (!diag_lane && row_ok) {
#pragma unroll
    for (int c = 0; c < SUB; ++c) trow[c * TSTR] = a[c];
  }
  return trow;
}

// FUSED: the nt workgroups of a supernode (one per TRSM row tile) each
// factor the block; the inverse tiles then replace the factored diagonal
// tiles in LDS (and 1/D of every column is kept for LDL).  The factored block
// and the inverse tiles are written to HBM by the workgroup that is the last
// of the nt to finish loading the block (*cnt counts the loads, and is reset
// by that workgroup): the others have read the unfactored block by then.
template <typename T, bool LDL, int NBM, bool FUSED>
__device__ __forceinline__ void diag_body(const Task4& t, const SnInfo& s, T* __restrict__ arena,
                                          T* __restrict__ dinv, int* __restrict__ info,
                                          DiagLds<T, NBM>& L, int* __restrict__ cnt, int nt) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NTL = NBM * (NBM + 1) / 2;
  T* __restrict__ tl = L.tl;
  T* __restrict__ dsh = L.dsh;
  T* __restrict__ ysh = L.ysh;
  T (*colb)[2][SUB] = L.colb;
  const int c0 = t.b * kStep;
  const int bs = min(kStep, s.w - c0);
  const int nb = (bs + SUB - 1) / SUB;
  const int64_t ld = s.r;
  T* __restrict__ P = arena + s.off + (int64_t)c0 * ld + c0;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar control

  // ---- load the lower tiles (clamped unconditional loads, one wait) ----
  {
    const int r = tid & 15, c = tid >> 4;
    T v[NTL];
#pragma unroll
    for (int I = 0; I < NBM; ++I)
#pragma unroll
      for (int J = 0; J <= I; ++J) {
        const int row = min(SUB * I + r, bs - 1), col = min(SUB * J + c, bs - 1);
        v[tslot(I, J)] = P[(int64_t)col * ld + row];
      }
#pragma unroll
    for (int q = 0; q < NTL; ++q) keep(v[q]);
#pragma unroll
    for (int I = 0; I < NBM; ++I)
#pragma unroll
      for (int J = 0; J <= I; ++J) {
        const int row = SUB * I + r, col = SUB * J + c;
        const bool in = row < bs && col < bs && row >= col;
        tl[tslot(I, J) * TSZ + c * TSTR + r] = in ? v[tslot(I, J)] : T(row == col ? 1 : 0);
      }
  }
  __syncthreads();
  if (FUSED && tid == 0) {
    // every load of this workgroup has returned (its values are in LDS)
    const int prior = atomicAdd(cnt, 1);
    L.store = (prior == nt - 1);
    if (prior == nt - 1) atomicExch(cnt, 0);
  }

  // Trailing tiles live in MFMA accumulators for the whole factorization:
  // tile t of the k-descending order belongs to wave t % 4, slot t / 4, and
  // goes back to LDS when its block column becomes the panel.
  constexpr int NS = ((NBM - 1) * NBM / 2 + 3) / 4;
  const int ntt = (nb - 1) * nb / 2;
  const unsigned char* __restrict__ ord = kOrd + kOrdOff[nb];
  v4 acc[NS > 0 ? NS : 1];
  int tI[NS > 0 ? NS : 1], tK[NS > 0 ? NS : 1];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const int t = 4 * q + wave;
    const int e = ord[min(t, max(ntt - 1, 0))];
    tI[q] = (t < ntt) ? (e >> 4) + 1 : 0;
    tK[q] = (t < ntt) ? (e & 15) + 1 : NBM;  // NBM: never active
    const T* C = tl + tslot(max(tI[q], tK[q] < NBM ? tK[q] : 0), tK[q] < NBM ? tK[q] : 0) * TSZ;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[q][v] = C[(lane & 15) * TSTR + Mfma<T>::row(lane, v)];
  }

  const bool diag_lane = lane < SUB;
  for (int J = 0; J < nb; ++J) {
    // ---- 1. block column J, lane per row (its tiles were written to LDS
    // by the initial load or at the end of step J-1) ----
    T a[SUB];
    T* trow = chain_column<T, LDL>(tl, colb[wave], J, nb, wave, lane, a);
    if (LDL && wave == 0 && diag_lane) dsh[lane] = a[lane];
    __syncthreads();
    // every wave has read the unfactored diagonal tile: wave 0 stores the
    // factored lower part (read again only after the loop)
    if (diag_lane && wave == 0) {
#pragma unroll
      for (int c = 0; c < SUB; ++c) trow[c * TSTR] = (c > lane) ? T(0) : a[c];
    }
    // ---- 2. trailing update A_IK -= L_IJ L_KJ^T of the register tiles with
    // K > J (a prefix of each wave's slots).  The operands of slot q+1 are
    // read from LDS before the four MFMAs of slot q issue, so their latency
    // hides behind the MFMA pipe ----
    const int m = nb - J - 1;
    const int ntr = m * (m + 1) / 2;
    const int na = (ntr - wave + 3) >> 2;  // active slots of this wave
    T dk[4];  // LDL: d_k of this lane's four k rows
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) dk[kk] = LDL ? dsh[Mfma<T>::row(lane, kk)] : T(1);
    // Loads are unconditional (an inactive slot reads some valid tile) and a
    // slot is reached only through the one before it, so the LDS wait
    // before slot q's MFMAs covers exactly slot q's operands.
    T av[2][4], bv[2][4];
    auto load_slot = [&](int q, T* ao, T* bo) {
      const T* A = tl + tslot(min(tI[q], nb - 1), J) * TSZ + (lane & 15);
      const T* B = tl + tslot(min(tK[q], nb - 1), J) * TSZ + (lane & 15);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int kp = Mfma<T>::row(lane, kk);
        ao[kk] = A[kp * TSTR];
        bo[kk] = B[kp * TSTR];
      }
    };
    auto load = [&](int q) { load_slot(q, av[q & 1], bv[q & 1]); };
    auto mma = [&](int q) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
        acc[q] = Mfma<T>::mma(LDL ? -(av[q & 1][kk] * dk[kk]) : -av[q & 1][kk], bv[q & 1][kk], acc[q]);
    };
    load(0);
    SlotChain<0, NS>::run(na, load, mma);
    // ---- 3. the tiles of block column J+1 are final: they leave the
    // accumulators (step 2 of other waves reads only column J) ----
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (J + 1 < nb && tK[q] == J + 1) {  // unused slots hold tK = NBM
        T* C = tl + tslot(tI[q], J + 1) * TSZ;
#pragma unroll
        for (int v = 0; v < 4; ++v) C[(lane & 15) * TSTR + Mfma<T>::row(lane, v)] = acc[q][v];
      }
    }
    __syncthreads();
  }

  // ---- failed pivots (a pivot that is not > 0, LL, or is 0 or NaN, LDL;
  // NaN spreads to every later pivot, so the smallest column is the first
  // failure) and, for LL, 1/L_jj for the inverse tiles ----
  if (tid < nb * SUB) {
    const int Jd = tid / SUB, jd = tid % SUB;
    const T dj = tl[tslot(Jd, Jd) * TSZ + jd * TSTR + jd];
    const bool bad = LDL ? (dj == T(0) || dj != dj) : !(dj > T(0));
    if (bad && tid < bs) atomicMin(info, s.fc + c0 + tid + 1);
    if (!LDL) ysh[tid] = T(1) / dj;
  }
  __syncthreads();

  // ---- inverses of the factored 16x16 diagonal tiles for the panel TRSM,
  // all at once: 16 lanes per tile, lane (c) solves L z = e_c (LDL: unit
  // diagonal) by forward substitution ----
  const bool store = FUSED ? (L.store != 0) : true;
  const int Jt = 4 * wave + (lane >> 4);
  const int cz = lane & 15;
  T z[SUB];
  if (Jt < nb) {
    const T* Lt = tl + tslot(Jt, Jt) * TSZ;
#pragma unroll
    for (int i = 0; i < SUB; ++i) {
      T acc = (cz == i) ? T(1) : T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) acc -= Lt[k * TSTR + i] * z[k];
      z[i] = LDL ? acc : acc * ysh[SUB * Jt + i];
    }
    if (store) {
      T* out = dinv + (int64_t)(t.c + Jt) * (SUB * SUB) + cz * SUB;
#pragma unroll
      for (int i = 0; i < SUB; ++i) out[i] = z[i];
    }
  }

  // ---- write the factored lower triangle back (LDS reads first, one wait) ----
  if (store) {
    const int r = tid & 15, c = tid >> 4;
    T v[NTL];
#pragma unroll
    for (int q = 0; q < NTL; ++q) v[q] = tl[q * TSZ + c * TSTR + r];
#pragma unroll
    for (int q = 0; q < NTL; ++q) keep(v[q]);
#pragma unroll
    for (int I = 0; I < NBM; ++I)
#pragma unroll
      for (int J = 0; J <= I; ++J) {
        const int row = SUB * I + r, col = SUB * J + c;
        if (row < bs && col < bs && row >= col) P[(int64_t)col * ld + row] = v[tslot(I, J)];
      }
  }

  if (FUSED) {
    // the TRSM reads the inverse tiles where the diagonal tiles were (and,
    // for LDL, 1/D from ysh): every read of the diagonal tiles above ends
    // before the first barrier, D is read before the second
    __syncthreads();
    if (LDL && tid < NBM * SUB) {
      const int Jd = tid / SUB, jd = tid % SUB;
      const T d = tl[tslot(Jd, Jd) * TSZ + jd * TSTR + jd];
      ysh[tid] = (tid < bs) ? T(1) / d : T(0);
    }
    __syncthreads();
    if (Jt < nb) {
#pragma unroll
      for (int i = 0; i < SUB; ++i) tl[tslot(Jt, Jt) * TSZ + cz * TSTR + i] = z[i];
    }
    __syncthreads();
  }
}

template <typename T, bool LDL, int NBM>
__global__ __launch_bounds__(NT) void k_diag(const SnInfo* __restrict__ sn,
                                              const Task4* __restrict__ tasks,
                                              T* __restrict__ arena, T* __restrict__ dinv,
                                              int* __restrict__ info) {
  __shared__ __attribute__((aligned(16))) DiagLds<T, NBM> L;
  const Task4 t = tasks[blockIdx.x];
  diag_body<T, LDL, NBM, false>(t, sn[t.a], arena, dinv, info, L, nullptr, 0);
}

// ---------------------------------------------------------------------------
// Panel TRSM of 64 rows below a factored diagonal block (blockCell_t::trsm,
// symPACKMatrix2D.hpp:1104-1178; LDL: blockCellLDL_t::trsm, :2118-2139):
// X = B L^{-T} (LDL: B L^{-T} D^{-1}, unit L), blocked by 16 columns with the
// inverted diagonal tiles of k_diag:
//   Y_K = inv(L_KK) (B_K^T - sum_{J<K} L_KJ Y_J),  X_K^T = Y_K (LDL: D_K^{-1} Y_K).
// Each wave owns 16 rows; Y^T lives in MFMA accumulators, and the f64/f32
// accumulator layout doubles as the B operand of the next products (slot kk
// of tile J is k-chunk kk), so no data moves between the steps.
// Task {sn, step, row_tile, inverse-tile offset}.
// ---------------------------------------------------------------------------
template <typename T, bool LDL, int NBM, int ROWS>
__device__ __forceinline__ void ptrsm_compute(const SnInfo& s, int c0, int bs, int nb, int rt,
                                              T* __restrict__ P, const T* __restrict__ tl,
                                              const T* __restrict__ dd);

// 128 rows per workgroup (8 waves): the LDS-resident block (up to 119 KB)
// allows one workgroup per CU, so each holds twice the rows of a 4-wave one
template <typename T, bool LDL, int NBM>
__global__ __launch_bounds__(2 * NT) void k_ptrsm(const SnInfo* __restrict__ sn,
                                               const Task4* __restrict__ tasks,
                                               T* __restrict__ arena, const T* __restrict__ dinv) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NTL = NBM * (NBM + 1) / 2;
  __shared__ __attribute__((aligned(16))) T tl[NTL * TSZ];  // (I,J) I>J: L; (I,I): inv(L_II)
  __shared__ T dd[NBM * SUB];                               // LDL: 1/D
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int c0 = t.b * kStep;
  const int bs = min(kStep, s.w - c0);
  const int nb = (bs + SUB - 1) / SUB;
  const int64_t ld = s.r;
  T* __restrict__ P = arena + s.off + (int64_t)c0 * ld;  // column c0 of the panel
  const int tid = threadIdx.x;
  if (tid < NT) {  // the first 256 threads stage the block
    const int r = tid & 15, c = tid >> 4;
    const T* Dv = dinv + (int64_t)t.d * (SUB * SUB);
    T v[NTL];
#pragma unroll
    for (int I = 0; I < NBM; ++I)
#pragma unroll
      for (int J = 0; J <= I; ++J) {
        if (I == J) {
          v[tslot(I, J)] = Dv[(int64_t)min(I, nb - 1) * (SUB * SUB) + c * SUB + r];
        } else {
          const int row = min(SUB * I + r, bs - 1), col = min(SUB * J + c, bs - 1);
          v[tslot(I, J)] = P[(int64_t)col * ld + c0 + row];
        }
      }
#pragma unroll
    for (int q = 0; q < NTL; ++q) keep(v[q]);
#pragma unroll
    for (int I = 0; I < NBM; ++I)
#pragma unroll
      for (int J = 0; J <= I; ++J) {
        const int row = SUB * I + r, col = SUB * J + c;
        const bool in = (I == J) ? (I < nb) : (row < bs && col < bs);
        tl[tslot(I, J) * TSZ + c * TSTR + r] = in ? v[tslot(I, J)] : T(0);
      }
    if (LDL && tid < NBM * SUB) {
      const int j = min(tid, bs - 1);
      const T d = P[(int64_t)j * ld + c0 + j];
      dd[tid] = (tid < bs) ? T(1) / d : T(0);
    }
  }
  __syncthreads();
  ptrsm_compute<T, LDL, NBM, kTrsmRowsWide>(s, c0, bs, nb, t.c, P, tl, dd);
}

// The TRSM proper: rows of row tile rt below the block, with the strict
// lower tiles and the inverse diagonal tiles in LDS (tl) and, for LDL, 1/D
// of every column (dd).  P points at column c0 of the panel.
template <typename T, bool LDL, int NBM, int ROWS>
__device__ __forceinline__ void ptrsm_compute(const SnInfo& s, int c0, int bs, int nb, int rt,
                                              T* __restrict__ P, const T* __restrict__ tl,
                                              const T* __restrict__ dd) {
  typedef typename Mfma<T>::v4 v4;
  const int64_t ld = s.r;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // rows of this wave: R0 + (lane & 15), columns as MFMA rows
  const int R0 = c0 + bs + rt * ROWS + 16 * wave;
  if (R0 >= s.r) return;  // no barrier follows
  const int prow = min(R0 + (lane & 15), s.r - 1);
  const bool prow_ok = R0 + (lane & 15) < s.r;
  v4 y[NBM];
#pragma unroll
  for (int K = 0; K < NBM; ++K)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = min(SUB * K + Mfma<T>::row(lane, v), bs - 1);
      y[K][v] = P[(int64_t)col * ld + prow];
    }
#pragma unroll
  for (int K = 0; K < NBM; ++K) asm volatile("" : "+v"(y[K]));
#pragma unroll
  for (int K = 0; K < NBM; ++K) {
    if (K < nb) {
      // every LDS operand of block row K (the tiles L_KJ, J < K, and the
      // inverse tile) is read before the row's first product, so the
      // products issue back to back instead of each waiting for its read
      T lv[NBM][4], li[4];
#pragma unroll
      for (int J = 0; J < K; ++J) {
        const T* L = tl + tslot(K, J) * TSZ + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) lv[J][kk] = L[Mfma<T>::row(lane, kk) * TSTR];
      }
      {
        const T* Li = tl + tslot(K, K) * TSZ + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) li[kk] = Li[Mfma<T>::row(lane, kk) * TSTR];
      }
      // two accumulators (even / odd J) halve the dependent MFMA chain
      v4 y2 = v4{0, 0, 0, 0};
#pragma unroll
      for (int J = 0; J < K; ++J) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          if (J & 1)
            y2 = Mfma<T>::mma(-lv[J][kk], y[J][kk], y2);
          else
            y[K] = Mfma<T>::mma(-lv[J][kk], y[J][kk], y[K]);
        }
      }
      y[K] += y2;
      v4 z = v4{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) z = Mfma<T>::mma(li[kk], y[K][kk], z);
      y[K] = z;
    }
  }
  if (!prow_ok) return;
  T* __restrict__ out = P + R0 + (lane & 15);
#pragma unroll
  for (int K = 0; K < NBM; ++K) {
    if (K < nb) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = SUB * K + Mfma<T>::row(lane, v);
        T x = y[K][v];
        if (LDL) x *= dd[min(col, NBM * SUB - 1)];
        if (col < bs) out[(int64_t)col * ld] = x;
      }
    }
  }
}

// Block factor and panel TRSM in one workgroup per (supernode, row tile) of
// a supernode of width <= kStep: every workgroup of a supernode factors the
// diagonal block itself (used on etree levels with few supernodes, where the
// GPU would otherwise idle between the two launches).  Task {sn, number of
// row tiles, row tile, inverse-tile offset}; cnt: one zeroed counter per
// supernode.
template <typename T, bool LDL, int NBM>
__global__ __launch_bounds__(NT) void k_diagtrsm(const SnInfo* __restrict__ sn,
                                                  const Task4* __restrict__ tasks,
                                                  T* __restrict__ arena, T* __restrict__ dinv,
                                                  int* __restrict__ info, int* __restrict__ cnt) {
  __shared__ __attribute__((aligned(16))) DiagLds<T, NBM> L;
  const Task4 tk = tasks[blockIdx.x];
  const Task4 t{tk.a, 0, tk.d, 0};  // the block factor's task layout {sn, step, inverse tile}
  const SnInfo s = sn[tk.a];
  diag_body<T, LDL, NBM, true>(t, s, arena, dinv, info, L, cnt + tk.a, tk.b);
  const int bs = min(kStep, s.w);
  ptrsm_compute<T, LDL, NBM, kTrsmRows>(s, 0, bs, (bs + SUB - 1) / SUB, tk.c, arena + s.off, L.tl, L.ysh);
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
                                                const int32_t* __restrict__ rel,
                                                const MergeInfo* __restrict__ merges,
                                                const int32_t* __restrict__ mmap, int64_t ntasks,
                                                int ksplit) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int KC = kSchurKC;
  constexpr int LDS_LD = TILE + 1;  // odd pitch: the 4 k rows a wave reads hit different banks
  __shared__ T As[KC][LDS_LD];
  __shared__ T Bs[KC][LDS_LD];
  __shared__ T Dk[KC];
  __shared__ int64_t sTc[TILE], sTm[TILE];  // MODE 1: scatter bases of the tile's output columns
  // grid-stride over (task, k part) pairs (the launches use one workgroup
  // per pair); part `part` of a task covers k chunks [q_lo, q_hi)
  for (int64_t gi = blockIdx.x; gi < ntasks * ksplit; gi += gridDim.x) {
  const int64_t task = gi / ksplit;
  const int part = (int)(gi - task * ksplit);
  const Task4 t = tasks[task];
  const SnInfo s = sn[t.a];
  T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  int kbeg, K, arow0, brow0, alim, blim, ti, tj;
  if (MODE == 0) {
    kbeg = t.b * kStep;
    K = min(kStep, s.w - kbeg);
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
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
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
  // MODE 1 with t.d > 0: after the K columns of this panel, K2 more k
  // columns from a merged child's panel, whose rows are gathered through
  // the row map (rows the child lacks are zero)
  int K2 = 0;
  const T* __restrict__ P2 = P;
  int64_t ld2 = 1;
  int arow2 = 0, brow2 = 0;
  bool arow2_ok = false, brow2_ok = false;
  if (MODE == 1 && t.d > 0) {
    const MergeInfo mg = merges[t.d - 1];
    K2 = mg.w;
    P2 = arena + mg.off;
    ld2 = mg.ld;
    const int ma = mmap[mg.moff + (arow - s.w)];
    const int mb = mmap[mg.moff + (brow - s.w)];
    arow2_ok = arow_ok && ma >= 0;
    brow2_ok = brow_ok && mb >= 0;
    arow2 = max(ma, 0);
    brow2 = max(mb, 0);
  }
  const int nq1 = (K + KC - 1) / KC;
  const int nq = nq1 + (K2 + KC - 1) / KC;
  const int q_lo = (int)((int64_t)part * nq / ksplit), q_hi = (int)((int64_t)(part + 1) * nq / ksplit);
  if (q_lo >= q_hi) continue;  // uniform over the workgroup
  // A chunk is loaded from clamped (valid) addresses into ra/rb with its
  // validity bits in mk; the bits are applied when the chunk is written to
  // LDS, a chunk later, so nothing consumes the loads before the MFMAs of
  // the current chunk have issued.  LDL: dv holds d_k of k column tid.
  T ra[NQ], rb[NQ], dv = T(0);
  unsigned mk = 0;
  auto fetch = [&](int qc) {
    unsigned m = 0;
    if (qc < nq1) {
      const int kc = qc * KC;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int kg = kc + lk + 4 * q;
        const T* col = P + (int64_t)(kbeg + min(kg, K - 1)) * ld;
        ra[q] = col[arow];
        rb[q] = col[brow];
        m |= (unsigned)((kg < K) & arow_ok) << q;
        m |= (unsigned)((kg < K) & brow_ok) << (q + NQ);
      }
      if (LDL) {
        const int kg = min(kc + (tid & (KC - 1)), K - 1);
        dv = P[(int64_t)(kbeg + kg) * ld + kbeg + kg];
        m |= (unsigned)(kc + (tid & (KC - 1)) < K) << (2 * NQ);
      }
    } else {
      const int kc = (qc - nq1) * KC;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int kg = kc + lk + 4 * q;
        const T* col = P2 + (int64_t)min(kg, K2 - 1) * ld2;
        ra[q] = col[arow2];
        rb[q] = col[brow2];
        m |= (unsigned)((kg < K2) & arow2_ok) << q;
        m |= (unsigned)((kg < K2) & brow2_ok) << (q + NQ);
      }
      if (LDL) {
        const int kg = min(kc + (tid & (KC - 1)), K2 - 1);
        dv = P2[(int64_t)kg * ld2 + kg];
        m |= (unsigned)(kc + (tid & (KC - 1)) < K2) << (2 * NQ);
      }
    }
    mk = m;
  };
  // On a diagonal tile (ti == tj) the wave holding rows b of the upper half
  // and columns a of the lower half (wb > wa) only produces entries a < b,
  // which are never stored: it skips its products.
  const bool idle = (ti == tj) && (wb > wa);
  // MODE 1: the scatter bases tcol / tmap of the tile's 64 output columns b
  // are loaded with the first chunk and kept in LDS, so the epilogue needs
  // one memory round trip (rel) instead of two, and that one is issued
  // before the last chunk's products
  int64_t tcv = 0, tmv = 0;
  const int m = s.r - s.w;
  if (MODE == 1 && tid < TILE) {
    const int bc = min(tj * TILE + tid, m - 1);
    tcv = tcol[s.upd + bc];
    tmv = tmap[s.upd + bc];
  }
  fetch(q_lo);
  auto stage = [&](int qc) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      As[lk + 4 * q][lrow] = ((mk >> q) & 1u) ? ra[q] : T(0);
      Bs[lk + 4 * q][lrow] = ((mk >> (q + NQ)) & 1u) ? rb[q] : T(0);
    }
    if (LDL && tid < KC) Dk[tid] = ((mk >> (2 * NQ)) & 1u) ? dv : T(0);
    if (MODE == 1 && qc == q_lo && tid < TILE) {
      sTc[tid] = tcv;
      sTm[tid] = tmv;
    }
    __syncthreads();
  };
  auto products = [&](int qc) {
    // k steps of 4 that hold valid columns in this chunk (the last chunk of
    // each part is partial: its zero-filled tail is not multiplied)
    const int krem = qc < nq1 ? K - qc * KC : K2 - (qc - nq1) * KC;
    const int nkk = idle ? 0 : min(KC / 4, (krem + 3) >> 2);
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      if (kk >= nkk) break;  // uniform over the wave
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
  };
  for (int qc = q_lo; qc < q_hi - 1; ++qc) {
    stage(qc);
    fetch(qc + 1);  // next chunk in flight during the MFMAs
    products(qc);
    __syncthreads();
  }
  // last chunk (peeled: the prefetch registers are free, the scatter
  // offsets take their place)
  stage(q_hi - 1);
  // MODE 1: scatter offsets of the 16 outputs of this lane, target column
  // base from LDS plus the relative row (rel is defined for rows a >= the
  // first row of b's target group: clamped, unconditional loads)
  int64_t off[2][4][2];
  if (MODE == 1) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int bl = wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
        const int bc = min(tj * TILE + bl, m - 1);
        const int64_t tc = sTc[bl], tmb = sTm[bl];
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) {
          const int aoff = ti * TILE + wa * 32 + aj * 16 + (lane & 15);
          const int ac = min(max(aoff, bc), m - 1);
          off[bi][v][aj] = tc + rel[tmb + ac];
        }
      }
  }
  products(q_hi - 1);

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
    // a deferred supernode scatters only the output columns in its parent
    const int bcut = t.d < 0 ? -t.d : m;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int boff = tj * TILE + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) {
          const int aoff = ti * TILE + wa * 32 + aj * 16 + (lane & 15);
          if (boff < bcut && aoff < m && aoff >= boff) atomic_add(arena + off[bi][v][aj], -acc[bi][aj][v]);
        }
      }
  }
  // a workgroup that takes another (task, part) rewrites As / Bs / sTc only
  // after every wave has finished reading them
  if (gi + gridDim.x < ntasks * ksplit) __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Schur update with 128 x 64 output tiles (rows a x columns b; the
// autotuner's wide candidate).  Same algorithm as k_update<T,1,LDL>: C =
// -P_a P_b^T (LDL: -P_a D P_b^T) over the supernode's k columns (and a merged
// child's, gathered through the row map), K in chunks of 32 staged through
// LDS with the next chunk in flight, scatter through tcol/tmap/rel with
// atomics.  Each wave owns 64 a-rows x 32 b-columns, 4 x 2 MFMA 16x16x4
// blocks: every k step reads 6 LDS operands for 8 products (4 for 4 in the
// 64 x 64 tile), and a tile reads 192 rows of operands for 8192 outputs
// (128 for 4096).  Task {sn, 128-row tile ti, 64-column tile tj, d} with
// 64 tj <= 128 ti + 127; d as in k_update.
// ---------------------------------------------------------------------------
template <typename T, bool LDL>
__global__ __launch_bounds__(NT) void k_schur_wide(const SnInfo* __restrict__ sn,
                                                    const Task4* __restrict__ tasks,
                                                    T* __restrict__ arena,
                                                    const int64_t* __restrict__ tcol,
                                                    const int64_t* __restrict__ tmap,
                                                    const int32_t* __restrict__ rel,
                                                    const MergeInfo* __restrict__ merges,
                                                    const int32_t* __restrict__ mmap, int64_t ntasks,
                                                    int ksplit) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int KC = kSchurKC;
  constexpr int TA = 2 * TILE, TB = TILE;
  constexpr int NQ = KC / 4;
  __shared__ T As[KC][TA + 1];  // odd pitches: the 4 k rows a wave reads hit different banks
  __shared__ T Bs[KC][TB + 1];
  __shared__ T Dk[KC];
  __shared__ int64_t sTc[TB], sTm[TB];
  for (int64_t gi = blockIdx.x; gi < ntasks * ksplit; gi += gridDim.x) {
    const int64_t task = gi / ksplit;
    const int part = (int)(gi - task * ksplit);
    const Task4 t = tasks[task];
    const SnInfo s = sn[t.a];
    const T* __restrict__ P = arena + s.off;
    const int64_t ld = s.r;
    const int K = s.w, m = s.r - s.w;
    const int ti = t.b, tj = t.c;
    const int arow0 = s.w + ti * TA, brow0 = s.w + tj * TB;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wb = wave >> 1, wa = wave & 1;
    v4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = v4{0, 0, 0, 0};
    // loader: a wave reads 64 consecutive rows of one k column; A rows in two
    // halves of 64
    const int lrow = tid & 63, lk = tid >> 6;
    bool aok[2];
    int arow[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      aok[h] = arow0 + h * 64 + lrow < s.r;
      arow[h] = min(arow0 + h * 64 + lrow, s.r - 1);
    }
    const bool bok = brow0 + lrow < s.r;
    const int brow = min(brow0 + lrow, s.r - 1);
    int K2 = 0;
    const T* __restrict__ P2 = P;
    int64_t ld2 = 1;
    int arow2[2] = {0, 0}, brow2 = 0;
    bool aok2[2] = {false, false}, bok2 = false;
    if (t.d > 0) {
      const MergeInfo mg = merges[t.d - 1];
      K2 = mg.w;
      P2 = arena + mg.off;
      ld2 = mg.ld;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int ma = mmap[mg.moff + (arow[h] - s.w)];
        aok2[h] = aok[h] && ma >= 0;
        arow2[h] = max(ma, 0);
      }
      const int mb = mmap[mg.moff + (brow - s.w)];
      bok2 = bok && mb >= 0;
      brow2 = max(mb, 0);
    }
    const int nq1 = (K + KC - 1) / KC;
    const int nq = nq1 + (K2 + KC - 1) / KC;
    const int q_lo = (int)((int64_t)part * nq / ksplit), q_hi = (int)((int64_t)(part + 1) * nq / ksplit);
    if (q_lo >= q_hi) continue;  // uniform over the workgroup
    // chunk registers: A halves, B; validity bits applied at the LDS write
    T ra[2][NQ], rb[NQ], dv = T(0);
    unsigned mk = 0;
    auto fetch = [&](int qc) {
      unsigned mm = 0;
      const bool own = qc < nq1;
      const int kc = (own ? qc : qc - nq1) * KC;
      const int Kp = own ? K : K2;
      const T* __restrict__ Pp = own ? P : P2;
      const int64_t ldp = own ? ld : ld2;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int kg = kc + lk + 4 * q;
        const T* col = Pp + (int64_t)min(kg, Kp - 1) * ldp;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          ra[h][q] = col[own ? arow[h] : arow2[h]];
          mm |= (unsigned)((kg < Kp) & (own ? aok[h] : aok2[h])) << (h * NQ + q);
        }
        rb[q] = col[own ? brow : brow2];
        mm |= (unsigned)((kg < Kp) & (own ? bok : bok2)) << (2 * NQ + q);
      }
      if (LDL) {
        const int kg = min(kc + (tid & (KC - 1)), Kp - 1);
        dv = Pp[(int64_t)kg * ldp + kg];
        mm |= (unsigned)(kc + (tid & (KC - 1)) < Kp) << (3 * NQ);
      }
      mk = mm;
    };
    // a wave whose 64 x 32 block lies wholly above the diagonal (every a < b)
    // skips its products
    const bool idle = ti * TA + wa * 64 + 63 < tj * TB + wb * 32;
    int64_t tcv = 0, tmv = 0;
    if (tid < TB) {
      const int bc = min(tj * TB + tid, m - 1);
      tcv = tcol[s.upd + bc];
      tmv = tmap[s.upd + bc];
    }
    fetch(q_lo);
    auto stage = [&](int qc) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int h = 0; h < 2; ++h) As[lk + 4 * q][h * 64 + lrow] = ((mk >> (h * NQ + q)) & 1u) ? ra[h][q] : T(0);
        Bs[lk + 4 * q][lrow] = ((mk >> (2 * NQ + q)) & 1u) ? rb[q] : T(0);
      }
      if (LDL && tid < KC) Dk[tid] = ((mk >> (3 * NQ)) & 1u) ? dv : T(0);
      if (qc == q_lo && tid < TB) {
        sTc[tid] = tcv;
        sTm[tid] = tmv;
      }
      __syncthreads();
    };
    auto products = [&](int qc) {
      const int krem = qc < nq1 ? K - qc * KC : K2 - (qc - nq1) * KC;
      const int nkk = idle ? 0 : min(KC / 4, (krem + 3) >> 2);
#pragma unroll
      for (int kk = 0; kk < KC / 4; ++kk) {
        if (kk >= nkk) break;  // uniform over the wave
        const int k = kk * 4 + (lane >> 4);
        T bf[2], af[4];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          bf[bi] = Bs[k][wb * 32 + bi * 16 + (lane & 15)];
          if (LDL) bf[bi] *= Dk[k];
        }
#pragma unroll
        for (int aj = 0; aj < 4; ++aj) af[aj] = As[k][wa * 64 + aj * 16 + (lane & 15)];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int aj = 0; aj < 4; ++aj) acc[bi][aj] = Mfma<T>::mma(bf[bi], af[aj], acc[bi][aj]);
      }
    };
    for (int qc = q_lo; qc < q_hi - 1; ++qc) {
      stage(qc);
      fetch(qc + 1);
      products(qc);
      __syncthreads();
    }
    stage(q_hi - 1);
    // scatter offsets: column base from LDS, relative row per output (rel is
    // defined for rows a >= the first row of b's target group: clamped,
    // unconditional loads), issued before the last chunk's products
    int64_t tcb[2][4];
    int32_t relv[2][4][4];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int bl = wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
        const int bc = min(tj * TB + bl, m - 1);
        tcb[bi][v] = sTc[bl];
        const int64_t tmb = sTm[bl];
#pragma unroll
        for (int aj = 0; aj < 4; ++aj) {
          const int aoff = ti * TA + wa * 64 + aj * 16 + (lane & 15);
          relv[bi][v][aj] = rel[tmb + min(max(aoff, bc), m - 1)];
        }
      }
    products(q_hi - 1);
    const int bcut = t.d < 0 ? -t.d : m;  // deferred supernode: its parent's columns only
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int boff = tj * TB + wb * 32 + bi * 16 + Mfma<T>::row(lane, v);
#pragma unroll
        for (int aj = 0; aj < 4; ++aj) {
          const int aoff = ti * TA + wa * 64 + aj * 16 + (lane & 15);
          if (boff < bcut && aoff < m && aoff >= boff)
            atomic_add(arena + tcb[bi][v] + relv[bi][v][aj], -acc[bi][aj][v]);
        }
      }
    if (gi + gridDim.x < ntasks * ksplit) __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Schur update with one wave per 64 x 64 output tile (the autotuner's
// candidate without LDS staging).  The same C = -P_a P_b^T (LDL:
// -P_a D P_b^T) as k_update<T,1,LDL>, but each wave owns a whole tile as 4 x 4
// MFMA 16x16x4 blocks (16 independent accumulators: one wave keeps the
// matrix core busy on its own) and streams its operands from memory straight
// into registers, two k steps of 4 ahead of the products, so there is no
// barrier: the four waves of a workgroup are four independent (task, part)
// items, and two waves per SIMD (about 200 registers each) overlap one
// tile's scatter with another's products.  Operand rows past the panel are
// clamped, not zeroed (they only reach outputs a >= m or b >= m, which are
// never stored); k past the width and a merged child's missing rows are
// zeroed.  A diagonal tile loads one row set for both sides and skips the
// six blocks above the diagonal; edge and deferred tiles skip the blocks
// past m or past the column cut.  Scatter as in k_update: column bases from
// a wave-private LDS copy of tcol / tmap, relative rows from rel, atomics.
// ---------------------------------------------------------------------------
template <typename T, bool LDL, int SHAPE>
__device__ __forceinline__ void w1_products(typename Mfma<T>::v4 (&acc)[4][4], const T (&ra)[4], const T (&rb)[4],
                                            T d, int nbi, int nbj) {
  // SHAPE 0: all 16 blocks; 1: diagonal tile, blocks with i >= j (rb = ra);
  // 2: blocks i < |nbi|, j < nbj, and i >= j when nbi < 0 (a diagonal tile)
  T bf[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bf[j] = LDL ? rb[j] * d : rb[j];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (SHAPE == 1 && i < j) continue;
      if (SHAPE == 2) {
        const bool dg = nbi < 0;
        const int ni = dg ? -nbi : nbi;
        if (i >= ni || j >= nbj || (dg && i < j)) continue;  // uniform
      }
      acc[j][i] = Mfma<T>::mma(bf[j], ra[i], acc[j][i]);
    }
}

// Loads one element through a buffer resource; offsets at or past the
// resource's size read 0.
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

// Per-wave body of k_schur_w1, specialized on the tile shape (SHAPE as in
// w1_products) and on whether the task carries a merged child's k columns.
// Operands come through buffer resources spanning the panel (and the merged
// child's panel): a k column past the width, a child's missing row (offset
// 0xffffffff) and, for LDL, d_k past the width all read 0, so the products
// need no masks.  Panels are below 2 GiB (the launch checks it).
template <typename T, bool LDL, int SHAPE, bool MERGED>
__device__ __forceinline__ void w1_body(const SnInfo& s, const Task4& t, T* __restrict__ arena,
                                        const int32_t* __restrict__ rel, const MergeInfo* __restrict__ merges,
                                        const int32_t* __restrict__ mmap, const int64_t* __restrict__ sTc,
                                        const int64_t* __restrict__ sTm, int g_lo, int g_hi, int ng1, int nbi,
                                        int nbj) {
  typedef typename Mfma<T>::v4 v4;
  constexpr uint32_t TB = sizeof(T);
  const int lane = threadIdx.x & 63;
  const uint32_t ld = (uint32_t)s.r;
  const int K = s.w, m = s.r - s.w, ti = t.b, tj = t.c;
  const int r16 = lane & 15, kq = lane >> 4;
  constexpr bool DIAG = SHAPE == 1;
  const __amdgpu_buffer_rsrc_t rP = make_rsrc(arena + s.off, (uint32_t)s.r * (uint32_t)s.w * TB);
  // byte offsets of this lane's operand rows in each 16-row block (clamped
  // into the panel: rows past m only reach outputs that are never stored)
  uint32_t aoff[4], boff[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    aoff[i] = (uint32_t)(s.w + min(ti * TILE + 16 * i + r16, m - 1)) * TB;
    boff[i] = (uint32_t)(s.w + min(tj * TILE + 16 * i + r16, m - 1)) * TB;
  }
  __amdgpu_buffer_rsrc_t rC = rP;
  uint32_t ld2 = 1;
  uint32_t aoff2[4] = {0, 0, 0, 0}, boff2[4] = {0, 0, 0, 0};
  if (MERGED) {
    const MergeInfo mg = merges[t.d - 1];
    rC = make_rsrc(arena + mg.off, (uint32_t)mg.ld * (uint32_t)mg.w * TB);
    ld2 = (uint32_t)mg.ld;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ma = mmap[mg.moff + (aoff[i] / TB - s.w)];
      const int mb = mmap[mg.moff + (boff[i] / TB - s.w)];
      aoff2[i] = ma >= 0 ? (uint32_t)ma * TB : 0xffffffffu;
      boff2[i] = mb >= 0 ? (uint32_t)mb * TB : 0xffffffffu;
    }
  }
  v4 acc[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[j][i] = v4{0, 0, 0, 0};

  // operand registers of one k step (4 k columns: lane l holds k = l / 16)
  struct Ops {
    T a[4], b[4], d;
  };
  // step g of this part; steps at or past g_hi read zeros (offsets past the
  // resource), so the loop below can always run whole pairs of steps
  auto fetch = [&](int g, Ops& o) {
    const bool child = MERGED && g >= ng1;
    const uint32_t k = (uint32_t)(4 * (child ? g - ng1 : g) + kq);
    const __amdgpu_buffer_rsrc_t r = child ? rC : rP;
    const uint32_t col = g < g_hi ? k * (child ? ld2 : ld) * TB : 0xffffffffu;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t ao = child ? aoff2[i] : aoff[i];
      o.a[i] = bload<T>(r, (ao == 0xffffffffu || col == 0xffffffffu) ? 0xffffffffu : col + ao);
    }
    if (!DIAG) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t bo = child ? boff2[i] : boff[i];
        o.b[i] = bload<T>(r, (bo == 0xffffffffu || col == 0xffffffffu) ? 0xffffffffu : col + bo);
      }
    }
    if (LDL) o.d = bload<T>(r, col == 0xffffffffu ? col : col + k * TB);
  };
  auto compute = [&](const Ops& o) {
    w1_products<T, LDL, SHAPE>(acc, o.a, DIAG ? o.a : o.b, LDL ? o.d : T(1), nbi, nbj);
  };
  // the k loop over triples of steps: steps g + 1 and g + 2 are in flight
  // during the products of step g (a count that is not a multiple of three
  // multiplies one or two steps of zeros); the scheduling barriers keep each
  // step's loads ahead of the products they overlap
  Ops o0, o1, o2;
  fetch(g_lo, o0);
  fetch(g_lo + 1, o1);
  for (int g = g_lo; g < g_hi; g += 3) {
    fetch(g + 2, o2);
    __builtin_amdgcn_sched_barrier(0);
    compute(o0);
    __builtin_amdgcn_sched_barrier(0);
    fetch(g + 3, o0);
    __builtin_amdgcn_sched_barrier(0);
    compute(o1);
    __builtin_amdgcn_sched_barrier(0);
    fetch(g + 4, o1);
    __builtin_amdgcn_sched_barrier(0);
    compute(o2);
    __builtin_amdgcn_sched_barrier(0);
  }

  // scatter: bases from LDS, relative rows from rel (defined for rows a >=
  // the first row of b's target group: clamped, unconditional loads), one
  // 16-column block at a time
  const int bcut = t.d < 0 ? -t.d : m;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int32_t rv[4][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int bl = 16 * j + Mfma<T>::row(lane, v);
      const int bc = min(tj * TILE + bl, m - 1);
      const int64_t tmb = sTm[bl];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ac = min(max(ti * TILE + 16 * i + r16, bc), m - 1);
        rv[v][i] = rel[tmb + ac];
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int bl = 16 * j + Mfma<T>::row(lane, v);
      const int b = tj * TILE + bl;
      const int64_t tc = sTc[bl];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int a = ti * TILE + 16 * i + r16;
        if (b < bcut && a < m && a >= b) atomic_add(arena + tc + rv[v][i], -acc[j][i][v]);
      }
    }
  }
}

template <typename T, bool LDL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_schur_w1(const SnInfo* __restrict__ sn, const Task4* __restrict__ tasks, T* __restrict__ arena,
                const int64_t* __restrict__ tcol, const int64_t* __restrict__ tmap,
                const int32_t* __restrict__ rel, const MergeInfo* __restrict__ merges,
                const int32_t* __restrict__ mmap, int64_t ntasks, int ksplit) {
  constexpr int KC = kSchurKC;
  __shared__ int64_t sTc[NT / 64][TILE], sTm[NT / 64][TILE];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t gi = (int64_t)blockIdx.x * (NT / 64) + wv;
  if (gi >= ntasks * ksplit) return;  // the whole wave
  const int64_t task = gi / ksplit;
  const int part = (int)(gi - task * ksplit);
  const Task4 t = tasks[task];
  const SnInfo s = sn[t.a];
  const int K = s.w, m = s.r - s.w, ti = t.b, tj = t.c;
  int K2 = 0;
  if (t.d > 0) K2 = merges[t.d - 1].w;
  // scatter bases of the tile's 64 output columns (wave-private LDS: the
  // wave's own LDS accesses complete in order, no barrier)
  {
    const int bc = min(tj * TILE + lane, m - 1);
    sTc[wv][lane] = tcol[s.upd + bc];
    sTm[wv][lane] = tmap[s.upd + bc];
  }
  // k runs in MFMA steps of 4; chunks of KC = 32 (8 steps) are the unit of
  // the k split, as in k_update
  constexpr int SPC = KC / 4;
  const int nq1 = (K + KC - 1) / KC, nq = nq1 + (K2 + KC - 1) / KC;
  const int q_lo = (int)((int64_t)part * nq / ksplit), q_hi = (int)((int64_t)(part + 1) * nq / ksplit);
  if (q_lo >= q_hi) return;
  const int ng1 = (K + 3) / 4, ng2 = (K2 + 3) / 4;
  const int g_lo = q_lo < nq1 ? SPC * q_lo : ng1 + SPC * (q_lo - nq1);
  const int g_hi = q_hi - 1 < nq1 ? min(SPC * q_hi, ng1) : min(ng1 + SPC * (q_hi - nq1), ng1 + ng2);
  // blocks of this tile that hold stored outputs
  const bool diag = ti == tj;
  const int nbi = min(4, (m - ti * TILE + 15) / 16);
  const int bcut = t.d < 0 ? -t.d : m;
  const int nbj = min(4, (min(bcut, m) - tj * TILE + 15) / 16);
  const int64_t* c = sTc[wv];
  const int64_t* mm = sTm[wv];
  if (nbi == 4 && nbj == 4 && !diag) {
    if (t.d > 0)
      w1_body<T, LDL, 0, true>(s, t, arena, rel, merges, mmap, c, mm, g_lo, g_hi, ng1, 4, 4);
    else
      w1_body<T, LDL, 0, false>(s, t, arena, rel, merges, mmap, c, mm, g_lo, g_hi, ng1, 4, 4);
  } else if (nbi == 4 && nbj == 4) {
    if (t.d > 0)
      w1_body<T, LDL, 1, true>(s, t, arena, rel, merges, mmap, c, mm, g_lo, g_hi, ng1, 4, 4);
    else
      w1_body<T, LDL, 1, false>(s, t, arena, rel, merges, mmap, c, mm, g_lo, g_hi, ng1, 4, 4);
  } else {
    const int nbi_enc = diag ? -nbi : nbi;
    if (t.d > 0)
      w1_body<T, LDL, 2, true>(s, t, arena, rel, merges, mmap, c, mm, g_lo, g_hi, ng1, nbi_enc, nbj);
    else
      w1_body<T, LDL, 2, false>(s, t, arena, rel, merges, mmap, c, mm, g_lo, g_hi, ng1, nbi_enc, nbj);
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

// Diagonal-block solve on MFMA (w <= 160), the panel-TRSM recurrence with
// right-hand sides in place of panel rows: for 16-row blocks t
//   forward   x_t = inv(L_tt) (y_t - sum_{u<t} L_tu x_u)
//   backward  x_t = inv(L_tt)^T (y_t - sum_{u>t} L_ut^T x_u)
// Each wave owns 16 right-hand sides (x^T in MFMA accumulators, whose
// layout is the B operand of the next product); a workgroup covers 64.
// LDL: unit inverse tiles, D applied by k_scale_d.  Task {sn, inverse-tile
// base, -, -}, blockIdx.y = chunk of 64 right-hand sides; Y may alias X.
template <typename T, bool BWD, int NBM>
__global__ __launch_bounds__(NT) void k_sdiag3(const SnInfo* __restrict__ sn,
                                               const Task4* __restrict__ tasks,
                                               const T* __restrict__ arena, const T* __restrict__ dinv,
                                               const T* Y, T* X, int nrhs) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NTL = NBM * (NBM + 1) / 2;
  __shared__ __attribute__((aligned(16))) T tl[NTL * TSZ];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int w = s.w;
  const int nbt = (w + SUB - 1) / SUB;
  const T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    const int r = tid & 15, c = tid >> 4;
    const T* Dv = dinv + (int64_t)t.b * (SUB * SUB);
    T v[NTL];
#pragma unroll
    for (int U = 0; U < NBM; ++U)
#pragma unroll
      for (int J = 0; J <= U; ++J) {
        if (U == J) {
          v[tslot(U, J)] = Dv[(int64_t)min(U, nbt - 1) * (SUB * SUB) + c * SUB + r];
        } else {
          const int row = min(SUB * U + r, w - 1), col = min(SUB * J + c, w - 1);
          v[tslot(U, J)] = P[(int64_t)col * ld + row];
        }
      }
#pragma unroll
    for (int q = 0; q < NTL; ++q) keep(v[q]);
#pragma unroll
    for (int U = 0; U < NBM; ++U)
#pragma unroll
      for (int J = 0; J <= U; ++J) {
        const int row = SUB * U + r, col = SUB * J + c;
        const bool in = (U == J) ? (U < nbt) : (row < w && col < w);
        tl[tslot(U, J) * TSZ + c * TSTR + r] = in ? v[tslot(U, J)] : T(0);
      }
  }
  __syncthreads();
  const int jb = blockIdx.y * 64 + 16 * wave;
  if (jb >= nrhs) return;  // no barrier follows
  const int j = jb + (lane & 15);
  const int jcl = min(j, nrhs - 1);
  const T jm = T(j < nrhs);
  const int64_t base = (int64_t)s.fc * nrhs + jcl;
  v4 x[NBM];
#pragma unroll
  for (int K = 0; K < NBM; ++K)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = min(SUB * K + Mfma<T>::row(lane, v), w - 1);
      x[K][v] = Y[base + (int64_t)row * nrhs] * jm;
    }
#pragma unroll
  for (int K = 0; K < NBM; ++K) asm volatile("" : "+v"(x[K]));
#pragma unroll
  for (int Ks = 0; Ks < NBM; ++Ks) {
    const int K = BWD ? NBM - 1 - Ks : Ks;
    if (K < nbt) {
      v4 x2 = v4{0, 0, 0, 0};
#pragma unroll
      for (int J = 0; J < NBM; ++J) {
        if (BWD ? (J > K && J < nbt) : (J < K)) {
          const T* L = BWD ? tl + tslot(J, K) * TSZ + (lane & 15) * TSTR  // L_JK^T[m][k'] = L_JK[k'][m]
                           : tl + tslot(K, J) * TSZ + (lane & 15);         // L_KJ[m][k']
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int kp = Mfma<T>::row(lane, kk);
            const T a = BWD ? L[kp] : L[kp * TSTR];
            if (J & 1)
              x2 = Mfma<T>::mma(-a, x[J][kk], x2);
            else
              x[K] = Mfma<T>::mma(-a, x[J][kk], x[K]);
          }
        }
      }
      x[K] += x2;
      const T* Li = BWD ? tl + tslot(K, K) * TSZ + (lane & 15) * TSTR : tl + tslot(K, K) * TSZ + (lane & 15);
      v4 z = v4{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int kp = Mfma<T>::row(lane, kk);
        z = Mfma<T>::mma(BWD ? Li[kp] : Li[kp * TSTR], x[K][kk], z);
      }
      x[K] = z;
    }
  }
  if (j >= nrhs) return;
#pragma unroll
  for (int K = 0; K < NBM; ++K) {
    if (K < nbt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = SUB * K + Mfma<T>::row(lane, v);
        if (row < w) X[base + (int64_t)row * nrhs] = x[K][v];
      }
    }
  }
}

// Diagonal-block solve for up to 16 right-hand sides (w <= 160), the
// 16-row blocks spread over the four waves (block K belongs to wave K % 4,
// its x^T in MFMA accumulators).  Each wave loads the tiles it needs (the
// L tiles that fold other blocks into its own, and its inverse diagonal
// tiles) straight from HBM into MFMA operand registers, all before the
// first product.  When block K is final its owner publishes it in LDS and
// every wave folds it into the blocks it owns that still wait for it
// (forward: K' > K; backward: K' < K), the next block first.  Same
// recurrence as k_sdiag3, which handles each right-hand-side group with one
// wave.  Task {sn, inverse-tile base, -, -}; Y may alias X.
template <typename T, bool BWD, int NBM>
__device__ __forceinline__ void sdiag4_body(const SnInfo& s, int64_t invb,
                                            const T* __restrict__ arena, const T* __restrict__ dinv,
                                            const T* Y, T* X, int nrhs, int j0, T (*xs)[4][64]) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NO = (NBM + 3) / 4;  // blocks per wave
  const int w = s.w;
  const int nbt = (w + SUB - 1) / SUB;
  const T* __restrict__ P = arena + s.off;
  const int64_t ld = s.r;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const T* __restrict__ Dv = dinv + invb * (SUB * SUB);
  const int m16 = lane & 15;
  // Loads in the order the recurrence consumes them (right-hand sides,
  // inverse tiles, then the fold operands block by block), so the first
  // steps wait only for the first loads.
  const int j = j0 + (lane & 15);
  const int jcl = min(j, nrhs - 1);
  const T jm = T(j < nrhs);
  const int64_t base = (int64_t)s.fc * nrhs + jcl;
  v4 xo[NO];
#pragma unroll
  for (int i = 0; i < NO; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = min(SUB * (wave + 4 * i) + Mfma<T>::row(lane, v), w - 1);
      xo[i][v] = Y[base + (int64_t)row * nrhs] * jm;
    }
  // Li[i][kk]: inverse diagonal tile of owned block i; Lf[i][K][kk] folds
  // block K into owned block i.  L(r, c) sits at P[c ld + r].
  T Lf[NO][NBM][4], Li[NO][4];
#pragma unroll
  for (int i = 0; i < NO; ++i)
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int kp = Mfma<T>::row(lane, kk);
      const int Kc = min(wave + 4 * i, nbt - 1);
      // forward uses inv(L_KK)(m16, kp), backward its transpose
      Li[i][kk] = Dv[(int64_t)Kc * (SUB * SUB) + (BWD ? m16 * SUB + kp : kp * SUB + m16)];
    }
#pragma unroll
  for (int Ks = 0; Ks < NBM; ++Ks) {
    const int K = BWD ? NBM - 1 - Ks : Ks;
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      const int K2 = wave + 4 * i;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int kp = Mfma<T>::row(lane, kk);
        // forward: L(16 K2 + m16, 16 K + kp); backward (L^T): L(16 K + kp, 16 K2 + m16)
        const int rr = BWD ? SUB * K + kp : SUB * K2 + m16;
        const int cc = BWD ? SUB * K2 + m16 : SUB * K + kp;
        // zero outside the block: the padding entries of x are not zero
        const bool need = K2 < nbt && K < nbt && (BWD ? K > K2 : K < K2) && rr < w && cc < w;
        Lf[i][K][kk] = need ? P[(int64_t)min(cc, w - 1) * ld + min(rr, w - 1)] : T(0);
      }
    }
  }
#pragma unroll
  for (int Ks = 0; Ks < NBM; ++Ks) {
    const int K = BWD ? NBM - 1 - Ks : Ks;
    if (K < nbt) {
      if ((K & 3) == wave) {
        // block K is complete: apply the inverse diagonal tile, publish
        const int i = K >> 2;
        v4 z = v4{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) z = Mfma<T>::mma(Li[i][kk], xo[i][kk], z);
        xo[i] = z;
#pragma unroll
        for (int v = 0; v < 4; ++v) xs[K][v][lane] = z[v];
      }
      __syncthreads();
      // fold block K into this wave's waiting blocks, the next one first
#pragma unroll
      for (int ii = 0; ii < NO; ++ii) {
        const int i = BWD ? NO - 1 - ii : ii;
        const int K2 = wave + 4 * i;
        if (K2 < nbt && (BWD ? K2 < K : K2 > K)) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) xo[i] = Mfma<T>::mma(-Lf[i][K][kk], xs[K][kk][lane], xo[i]);
        }
      }
    }
  }
  if (j < nrhs) {
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      const int K = wave + 4 * i;
      if (K < nbt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = SUB * K + Mfma<T>::row(lane, v);
          if (row < w) X[base + (int64_t)row * nrhs] = xo[i][v];
        }
      }
    }
  }
  __syncthreads();  // xs is reused by the next call
}

template <typename T, bool BWD, int NBM>
__global__ __launch_bounds__(NT) void k_sdiag4(const SnInfo* __restrict__ sn,
                                               const Task4* __restrict__ tasks,
                                               const T* __restrict__ arena, const T* __restrict__ dinv,
                                               const T* Y, T* X, int nrhs) {
  __shared__ __attribute__((aligned(16))) T xs[NBM][4][64];
  const Task4 t = tasks[blockIdx.x];
  sdiag4_body<T, BWD, NBM>(sn[t.a], t.b, arena, dinv, Y, X, nrhs, 0, xs);
}

// Multi-RHS off-diagonal solve updates on MFMA (used for nrhs >= 4), one
// workgroup per (64 panel rows below the diagonal block, 16 right-hand
// sides); wave q owns rows 16q..16q+15.  Task {sn, row tile, -, -},
// blockIdx.y = right-hand-side chunk.
// Forward:  Y[row(a)] -= sum_k L(a,k) X_I(k)      (one atomic per output)
template <typename T>
__device__ __forceinline__ void soffm_fwd_body(const SnInfo& s, int rt, const T* __restrict__ arena,
                                               const int32_t* __restrict__ lindx0,
                                               const T* __restrict__ X, T* __restrict__ Y, int nrhs,
                                               int j0) {
  typedef typename Mfma<T>::v4 v4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = s.r - s.w;
  const int a0 = rt * TILE + 16 * wave;
  if (a0 >= m) return;
  const int mi = lane & 15, g = lane >> 4;
  const int ac = min(a0 + mi, m - 1);
  const T amask = T(a0 + mi < m);
  const int jn = j0 + mi;
  const int jc = min(jn, nrhs - 1);
  const T jmask = T(jn < nrhs);
  const T* __restrict__ Pa = arena + s.off + s.w + ac;      // L(a, k) at Pa[k * ld]
  const T* __restrict__ Xk = X + (int64_t)s.fc * nrhs + jc;  // X_I(k, j) at Xk[k * nrhs]
  const int64_t ld = s.r;
  const int w = s.w;
  v4 acc = v4{0, 0, 0, 0};
  for (int k0 = 0; k0 < w; k0 += 32) {
    T av[8], bv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + 4 * q + g;
      const int kc = min(k, w - 1);
      const T km = T(k < w);
      av[q] = Pa[(int64_t)kc * ld] * (amask * km);
      bv[q] = Xk[(int64_t)kc * nrhs] * jmask;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) acc = Mfma<T>::mma(av[q], bv[q], acc);
  }
  if (jn >= nrhs) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int a = a0 + Mfma<T>::row(lane, v);
    if (a < m) atomic_add(Y + (int64_t)lindx0[s.lx + s.w + a] * nrhs + jn, -acc[v]);
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_soffm_fwd(const SnInfo* __restrict__ sn,
                                                   const Task4* __restrict__ tasks,
                                                   const T* __restrict__ arena,
                                                   const int32_t* __restrict__ lindx0,
                                                   const T* __restrict__ X, T* __restrict__ Y,
                                                   int nrhs) {
  const Task4 t = tasks[blockIdx.x];
  soffm_fwd_body<T>(sn[t.a], t.b, arena, lindx0, X, Y, nrhs, blockIdx.y * SUB);
}

// Backward:  X_I(k) -= sum_{a in tile} L(a,k) Y[row(a)]   (partial sums over
// 64 rows, atomics into X_I).  The reduction index is permuted so that lane
// (m, g) reads rows 16g..16g+15 of one column: one cache line per lane.
template <typename T>
__device__ __forceinline__ void soffm_bwd_body(const SnInfo& s, int rt, const T* __restrict__ arena,
                                               const int32_t* __restrict__ lindx0, T* __restrict__ X,
                                               int nrhs, int j0) {
  typedef typename Mfma<T>::v4 v4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = s.r - s.w;
  const int a0 = rt * TILE;
  const int mi = lane & 15, g = lane >> 4;
  const int jn = j0 + mi;
  const int jc = min(jn, nrhs - 1);
  const T jmask = T(jn < nrhs);
  const int64_t ld = s.r;
  // B operand (shared by every column block): Y rows of this tile
  int rowi[16];
#pragma unroll
  for (int st = 0; st < 16; ++st) rowi[st] = lindx0[s.lx + s.w + min(a0 + 16 * g + st, m - 1)];
  T bv[16];
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int a = a0 + 16 * g + st;
    bv[st] = X[(int64_t)rowi[st] * nrhs + jc] * (jmask * T(a < m));
  }
  // the column blocks of this wave (up to three per pass, w <= 192): every
  // load issued before the first MFMA
  const int nkb = (s.w + 15) >> 4;
  constexpr int MKB = 3;
  for (int kb0 = wave; kb0 < nkb; kb0 += MKB * (NT / 64)) {
    T av[MKB][16];
#pragma unroll
    for (int i = 0; i < MKB; ++i) {
      const int kcol = min((kb0 + i * (NT / 64)) * 16 + mi, s.w - 1);
      const T* __restrict__ Pc = arena + s.off + (int64_t)kcol * ld + s.w;
#pragma unroll
      for (int st = 0; st < 16; ++st) av[i][st] = Pc[min(a0 + 16 * g + st, m - 1)];
    }
#pragma unroll
    for (int i = 0; i < MKB; ++i)
#pragma unroll
      for (int st = 0; st < 16; ++st) keep(av[i][st]);
#pragma unroll
    for (int i = 0; i < MKB; ++i) {
      const int kb = kb0 + i * (NT / 64);
      if (kb < nkb) {
        v4 acc = v4{0, 0, 0, 0};
#pragma unroll
        for (int st = 0; st < 16; ++st) acc = Mfma<T>::mma(av[i][st], bv[st], acc);
        if (jn < nrhs) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int k = kb * 16 + Mfma<T>::row(lane, v);
            if (k < s.w) atomic_add(X + (int64_t)(s.fc + k) * nrhs + jn, -acc[v]);
          }
        }
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_soffm_bwd(const SnInfo* __restrict__ sn,
                                                   const Task4* __restrict__ tasks,
                                                   const T* __restrict__ arena,
                                                   const int32_t* __restrict__ lindx0,
                                                   T* __restrict__ X, int nrhs) {
  const Task4 t = tasks[blockIdx.x];
  soffm_bwd_body<T>(sn[t.a], t.b, arena, lindx0, X, nrhs, blockIdx.y * SUB);
}

// ---------------------------------------------------------------------------
// Partitioned-inverse solve (supernodes at most kStep wide).  Per supernode
// I (panel r x w, column-major, ld r) the solve factor
//   G_I = [I_w; L_off] L_II^{-1}
// (LDL: L_II unit lower) is kept in an arena of the panel's shape.  Then
//   forward  [X_I; dY] = G_I Y_I,  Y[struct rows] -= dY   (LDL: X_I /= D)
//   backward X_I = G_I^T [X_I; -X[struct rows]]
// (X_I = L_II^{-T}(X_I - L_off^T X_R) = L_II^{-T} X_I - (L_off L_II^{-1})^T X_R),
// one launch per etree level and sweep, no triangular recurrence on the
// solve's critical path.  Explicit inverses of the <= 160-wide diagonal
// blocks, as in the partitioned-inverse / selective-inversion methods
// (Alvarado, Pothen, Schreiber; Raghavan); the reference substitutes
// instead (forward_update / back_update, symPACKMatrix2D.hpp:1486-1797).
// ---------------------------------------------------------------------------

// Stage the strict lower tiles of the diagonal block (I > J) and the
// inverse diagonal tiles (I == J) of supernode s in LDS (all 256 threads).
template <typename T, int NBM>
__device__ __forceinline__ void stage_block_and_inverses(const SnInfo& s, int64_t invb,
                                                         const T* __restrict__ arena,
                                                         const T* __restrict__ dinv, T* __restrict__ tl) {
  constexpr int NTL = NBM * (NBM + 1) / 2;
  const int w = s.w;
  const int nb = (w + SUB - 1) / SUB;
  const int64_t ld = s.r;
  const T* __restrict__ P = arena + s.off;
  const int tid = threadIdx.x;
  const int r = tid & 15, c = tid >> 4;
  const T* Dv = dinv + invb * (SUB * SUB);
  T v[NTL];
#pragma unroll
  for (int I = 0; I < NBM; ++I)
#pragma unroll
    for (int J = 0; J <= I; ++J) {
      if (I == J) {
        v[tslot(I, J)] = Dv[(int64_t)min(I, nb - 1) * (SUB * SUB) + c * SUB + r];
      } else {
        const int row = min(SUB * I + r, w - 1), col = min(SUB * J + c, w - 1);
        v[tslot(I, J)] = P[(int64_t)col * ld + row];
      }
    }
#pragma unroll
  for (int q = 0; q < NTL; ++q) keep(v[q]);
#pragma unroll
  for (int I = 0; I < NBM; ++I)
#pragma unroll
    for (int J = 0; J <= I; ++J) {
      const int row = SUB * I + r, col = SUB * J + c;
      const bool in = (I == J) ? (I < nb) : (row < w && col < w);
      tl[tslot(I, J) * TSZ + c * TSTR + r] = in ? v[tslot(I, J)] : T(0);
    }
}

// G rows [64 rt, 64 rt + 64) of supernode t.a: X L_II = B with B = e_a^T for
// the diagonal rows a < w and B = L(a, :) below, i.e. the block backward
// substitution Y_K = inv(L_KK)^T (B_K^T - sum_{J>K} L_JK^T Y_J) on Y = X^T,
// 16 rows per wave (Y^T in MFMA accumulators, as in the panel TRSM).  Task
// {sn, row tile, inverse-tile base, -}.
template <typename T, int NBM>
__global__ __launch_bounds__(NT) void k_gprep(const SnInfo* __restrict__ sn, const Task4* __restrict__ tasks,
                                              const T* __restrict__ arena, const T* __restrict__ dinv,
                                              T* __restrict__ garena) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NTL = NBM * (NBM + 1) / 2;
  __shared__ __attribute__((aligned(16))) T tl[NTL * TSZ];
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  stage_block_and_inverses<T, NBM>(s, t.c, arena, dinv, tl);
  __syncthreads();
  const int w = s.w;
  const int nb = (w + SUB - 1) / SUB;
  const int64_t ld = s.r;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int R0 = t.b * TILE + 16 * wave;
  if (R0 >= s.r) return;  // no barrier follows
  const int prow = min(R0 + (lane & 15), s.r - 1);
  const bool prow_ok = R0 + (lane & 15) < s.r;
  const T* __restrict__ P = arena + s.off;
  v4 y[NBM];
#pragma unroll
  for (int K = 0; K < NBM; ++K)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = SUB * K + Mfma<T>::row(lane, v);
      const T lv = P[(int64_t)min(col, w - 1) * ld + prow];
      y[K][v] = col >= w ? T(0) : (prow < w ? T(prow == col ? 1 : 0) : lv);
    }
#pragma unroll
  for (int K = 0; K < NBM; ++K) asm volatile("" : "+v"(y[K]));
#pragma unroll
  for (int Ks = 0; Ks < NBM; ++Ks) {
    const int K = NBM - 1 - Ks;
    if (K < nb) {
      // the LDS operands of step K (L_JK^T for J > K, inv(L_KK)^T) are read
      // before its first product (clamped reads for J >= nb: not used)
      T lv[NBM][4], li[4];
#pragma unroll
      for (int J = K + 1; J < NBM; ++J) {
        const T* L = tl + tslot(min(J, nb - 1), K) * TSZ + (lane & 15) * TSTR;  // L_JK^T[m][k] = L_JK[k][m]
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) lv[J][kk] = L[Mfma<T>::row(lane, kk)];
      }
      {
        const T* Li = tl + tslot(K, K) * TSZ + (lane & 15) * TSTR;  // inv(L_KK)^T
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) li[kk] = Li[Mfma<T>::row(lane, kk)];
      }
      v4 y2 = v4{0, 0, 0, 0};
#pragma unroll
      for (int J = K + 1; J < NBM; ++J) {
        if (J < nb) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            if (J & 1)
              y2 = Mfma<T>::mma(-lv[J][kk], y[J][kk], y2);
            else
              y[K] = Mfma<T>::mma(-lv[J][kk], y[J][kk], y[K]);
          }
        }
      }
      y[K] += y2;
      v4 z = v4{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) z = Mfma<T>::mma(li[kk], y[K][kk], z);
      y[K] = z;
    }
  }
  if (!prow_ok) return;
  T* __restrict__ G = garena + s.off + prow;
#pragma unroll
  for (int K = 0; K < NBM; ++K) {
    if (K < nb) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = SUB * K + Mfma<T>::row(lane, v);
        if (col < w) G[(int64_t)col * ld] = y[K][v];
      }
    }
  }
}

// Forward level: rows [64 rt, 64 rt + 64) of [X_I; dY] = G_I Y_I, wave q
// owning 16 rows.  The right-hand sides go in blocks of 16: blocks
// blockIdx.y, blockIdx.y + gridDim.y, ... (a level with few tiles spreads
// the blocks over the grid; one with many walks them in one workgroup, so
// the G tile is read from HBM once and stays in registers).  Every G and
// Y_I operand of a block is loaded before its first product.  Rows a < w go
// to Xf (LDL: divided by D_a), rows below are subtracted from Y at their
// structure rows with atomics.  Y_I of a level's supernodes is never the
// target of the same level's updates (targets are ancestors).
template <typename T, bool LDL, int NBM>
__global__ __launch_bounds__(NT) void k_gsolve_fwd(const SnInfo* __restrict__ sn, const Task4* __restrict__ tasks,
                                                   const T* __restrict__ garena, const T* __restrict__ arena,
                                                   const int32_t* __restrict__ lindx0, T* Y, T* __restrict__ Xf,
                                                   int nrhs) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NQ = 4 * NBM;  // k chunks of 4
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int a0 = t.b * TILE + 16 * wave;
  if (a0 >= s.r) return;
  const int w = s.w;
  const int64_t ld = s.r;
  const int mi = lane & 15, g = lane >> 4;
  const int ac = min(a0 + mi, s.r - 1);
  const T* __restrict__ Ga = garena + s.off + ac;  // G(a, k) at Ga[k ld]
  T av[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int k = 4 * q + g;
    av[q] = Ga[(int64_t)min(k, w - 1) * ld] * T(k < w);
  }
  // output rows of this lane: structure rows below the block, or -1
  int orow[4];
  T dinv[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int a = a0 + Mfma<T>::row(lane, v);
    orow[v] = (a >= w && a < s.r) ? lindx0[s.lx + a] : -1;
    dinv[v] = T(1);
    if (LDL && a < w) dinv[v] = T(1) / arena[s.off + (int64_t)a * ld + a];
  }
  const int nblk = (nrhs + SUB - 1) / SUB;
  for (int jb = blockIdx.y; jb < nblk; jb += gridDim.y) {
    const int jn = jb * SUB + mi;
    const int jc = min(jn, nrhs - 1);
    const T jmask = T(jn < nrhs);
    const T* __restrict__ Yk = Y + (int64_t)s.fc * nrhs + jc;  // Y_I(k, j) at Yk[k nrhs]
    T bv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) bv[q] = Yk[(int64_t)min(4 * q + g, w - 1) * nrhs];
#pragma unroll
    for (int q = 0; q < NQ; ++q) keep(bv[q]);
    v4 acc = v4{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (4 * q < w)  // uniform: chunks wholly beyond w are skipped
        acc = Mfma<T>::mma(av[q], bv[q] * jmask, acc);
    }
    if (jn < nrhs) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int a = a0 + Mfma<T>::row(lane, v);
        if (a < w)
          Xf[(int64_t)(s.fc + a) * nrhs + jn] = LDL ? acc[v] * dinv[v] : acc[v];
        else if (orow[v] >= 0)
          atomic_add(Y + (int64_t)orow[v] * nrhs + jn, -acc[v]);
      }
    }
  }
}

// Backward level: X_I += G_I(tile rows, :)^T V(tile rows), V(a) = Xf(a) on
// the diagonal rows and -X(struct row a) below (the ancestors' final
// values).  Lane (m, g) reads rows 16g..16g+15 of one G column (a cache
// line), wave q the 16-column blocks q, q+4, q+8; the G values stay in
// registers while the workgroup walks its right-hand-side blocks (as in the
// forward level); every load of a block is issued before its first product;
// partial sums of the tile go to X_I with atomics (X zeroed before the
// sweep).
template <typename T, int NBM>
__global__ __launch_bounds__(NT) void k_gsolve_bwd(const SnInfo* __restrict__ sn, const Task4* __restrict__ tasks,
                                                   const T* __restrict__ garena,
                                                   const int32_t* __restrict__ lindx0, const T* __restrict__ Xf,
                                                   T* X, int nrhs) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int MKB = (NBM + 3) / 4;  // column blocks per wave
  const Task4 t = tasks[blockIdx.x];
  const SnInfo s = sn[t.a];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int w = s.w;
  const int nkb = (w + 15) >> 4;
  if (wave >= nkb) return;
  const int a0 = t.b * TILE;
  const int mi = lane & 15, g = lane >> 4;
  int rowi[16];
#pragma unroll
  for (int st = 0; st < 16; ++st) rowi[st] = lindx0[s.lx + min(a0 + 16 * g + st, s.r - 1)];
  T av[MKB][16];
#pragma unroll
  for (int i = 0; i < MKB; ++i) {
    const int kcol = min((wave + 4 * i) * 16 + mi, w - 1);
    const T* __restrict__ Gc = garena + s.off + (int64_t)kcol * s.r;
#pragma unroll
    for (int st = 0; st < 16; ++st) av[i][st] = Gc[min(a0 + 16 * g + st, s.r - 1)];
  }
  const int nblk = (nrhs + SUB - 1) / SUB;
  for (int jb = blockIdx.y; jb < nblk; jb += gridDim.y) {
    const int jn = jb * SUB + mi;
    const int jc = min(jn, nrhs - 1);
    const T jmask = T(jn < nrhs);
    T bv[16];
#pragma unroll
    for (int st = 0; st < 16; ++st) {
      const int a = a0 + 16 * g + st;
      const T* src = a < w ? Xf : X;
      const T sg = a < w ? T(1) : T(-1);
      bv[st] = src[(int64_t)rowi[st] * nrhs + jc] * (jmask * T(a < s.r) * sg);
    }
#pragma unroll
    for (int st = 0; st < 16; ++st) keep(bv[st]);
#pragma unroll
    for (int i = 0; i < MKB; ++i) {
      const int kb = wave + 4 * i;
      if (kb < nkb) {
        v4 acc = v4{0, 0, 0, 0};
#pragma unroll
        for (int st = 0; st < 16; ++st) acc = Mfma<T>::mma(av[i][st], bv[st], acc);
        if (jn < nrhs) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int k = kb * 16 + Mfma<T>::row(lane, v);
            if (k < w) atomic_add(X + (int64_t)(s.fc + k) * nrhs + jn, acc[v]);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dependency-driven solve (one persistent launch per sweep, no level
// barriers).  Workgroups take tasks in a topological order from a global
// ticket counter; a task waits only for the tasks it depends on, which hold
// smaller tickets and so were taken by workgroups already running: every
// wait ends.  Waits are bounded all the same: a wait that exceeds its
// budget raises the shared abort flag, after which every wait returns at
// once and the host reports the error (the result is then not used).
//
// Forward, supernodes in postorder: D(s) = diagonal solve of s, waiting for
// dcnt[s] = 0 (the off-diagonal tasks that update rows of s); then O(s, rt)
// for each 64-row tile of s, waiting for D(s), which subtract L_rt X_s from
// Y and count down dcnt of every supernode owning rows of the tile.
// Backward, supernodes in reverse postorder: R(s, rt) waits until every
// supernode owning rows of the tile is solved (bflag), subtracts
// L_rt^T X[rows] from X_s and counts down rcnt[s]; then B(s) waits for
// rcnt[s] = 0 and solves the diagonal block.
// The L2s of the XCDs are not coherent with each other: data written before
// a signal is released at agent scope (L2 write-back) and acquired by the
// waiter (L2 invalidate); the counters, flags and tickets themselves are
// read and updated with system-scope atomics, which bypass the XCD caches (a
// device-scope load may be served by the waiter's own L2 and never see
// another XCD's update).  Each wait is bounded in time (1 s of the 100 MHz
// real-time counter).
// Task {kind (0 D / 1 O / 2 R / 3 B), sn, row tile, first target}, targets
// tgt[first .. tend[task]).
// ---------------------------------------------------------------------------
struct DfState {
  int* head;     // [2] ticket counters (forward, backward)
  int* dcnt;     // per supernode: forward updates still to come
  int* dflag;    // per supernode: forward diagonal solve done
  int* rcnt;     // per supernode: backward row tiles still to come
  int* bflag;    // per supernode: backward diagonal solve done
  int* abort_;   // set when a wait exceeded its budget
};

__device__ __forceinline__ int ld_sys(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int add_sys(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// thread 0 waits until *p == want (or abort); every thread then acquires
__device__ __forceinline__ void df_wait(const int* p, int want, int* abort_) {
  if (threadIdx.x == 0) {
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    while (ld_sys(p) != want) {
      if (ld_sys(abort_)) break;
      __builtin_amdgcn_s_sleep(8);
      if (__builtin_amdgcn_s_memrealtime() - t0 > 100000000ull) {
        st_sys(abort_, 1);
        break;
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// every thread releases its writes, then thread 0 signals
__device__ __forceinline__ void df_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
}

__device__ __forceinline__ int df_ticket(int* head, int* slot) {
  if (threadIdx.x == 0) *slot = add_sys(head, 1);
  __syncthreads();
  const int t = *slot;
  __syncthreads();
  return t;
}

// counters for one solve (the abort flag is cleared by the host per Solve)
__global__ void k_df_reset(DfState st, const int* __restrict__ dcnt0, const int* __restrict__ rcnt0,
                           int ns) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
    st.dcnt[i] = dcnt0[i];
    st.rcnt[i] = rcnt0[i];
    st.dflag[i] = 0;
    st.bflag[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) st.head[threadIdx.x] = 0;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_fsolve_df(const SnInfo* __restrict__ sn,
                                                  const Task4* __restrict__ tasks,
                                                  const int32_t* __restrict__ tend, int ntasks,
                                                  const int32_t* __restrict__ tgt,
                                                  const T* __restrict__ arena, const T* __restrict__ dinv,
                                                  const int32_t* __restrict__ lindx0,
                                                  const int64_t* __restrict__ invb, T* Y, T* X,
                                                  int nrhs, DfState st) {
  constexpr int NBM = 10;
  __shared__ __attribute__((aligned(16))) T xs[NBM][4][64];
  __shared__ int slot;
  for (;;) {
    const int ti = df_ticket(st.head, &slot);
    if (ti >= ntasks) break;
    const Task4 t = tasks[ti];
    const SnInfo s = sn[t.b];
    if (t.a == 0) {
      df_wait(st.dcnt + t.b, 0, st.abort_);
      for (int j0 = 0; j0 < nrhs; j0 += SUB)
        sdiag4_body<T, false, NBM>(s, invb[t.b], arena, dinv, Y, X, nrhs, j0, xs);
      df_release();
      if (threadIdx.x == 0) st_sys(st.dflag + t.b, 1);
    } else {
      df_wait(st.dflag + t.b, 1, st.abort_);
      for (int j0 = 0; j0 < nrhs; j0 += SUB) soffm_fwd_body<T>(s, t.c, arena, lindx0, X, Y, nrhs, j0);
      df_release();
      if (threadIdx.x == 0)
        for (int q = t.d; q < tend[ti]; ++q) add_sys(st.dcnt + tgt[q], -1);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_bsolve_df(const SnInfo* __restrict__ sn,
                                                  const Task4* __restrict__ tasks,
                                                  const int32_t* __restrict__ tend, int ntasks,
                                                  const int32_t* __restrict__ tgt,
                                                  const T* __restrict__ arena, const T* __restrict__ dinv,
                                                  const int32_t* __restrict__ lindx0,
                                                  const int64_t* __restrict__ invb, T* X, int nrhs,
                                                  DfState st) {
  constexpr int NBM = 10;
  __shared__ __attribute__((aligned(16))) T xs[NBM][4][64];
  __shared__ int slot;
  for (;;) {
    const int ti = df_ticket(st.head + 1, &slot);
    if (ti >= ntasks) break;
    const Task4 t = tasks[ti];
    const SnInfo s = sn[t.b];
    if (t.a == 2) {
      for (int q = t.d; q < tend[ti]; ++q) df_wait(st.bflag + tgt[q], 1, st.abort_);
      for (int j0 = 0; j0 < nrhs; j0 += SUB) soffm_bwd_body<T>(s, t.c, arena, lindx0, X, nrhs, j0);
      df_release();
      if (threadIdx.x == 0) add_sys(st.rcnt + t.b, -1);
    } else {
      df_wait(st.rcnt + t.b, 0, st.abort_);
      for (int j0 = 0; j0 < nrhs; j0 += SUB)
        sdiag4_body<T, true, NBM>(s, invb[t.b], arena, dinv, X, X, nrhs, j0, xs);
      df_release();
      if (threadIdx.x == 0) st_sys(st.bflag + t.b, 1);
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

// b (column-major, original numbering) -> Y (row-major, permuted numbering).
// One workgroup (4 waves) per 64 consecutive original rows, through an LDS
// tile of 64 rows x 64 right-hand sides: a wave reads one column of b over
// the 64 rows (contiguous), the workgroup then writes each row's right-hand
// sides to its permuted row of Y (contiguous).  invp0[old] = new.
// Grid: ceil(n / 64).
template <typename T>
__global__ __launch_bounds__(256) void k_gather_rhs(const double* __restrict__ b,
                                                    T* __restrict__ Y,
                                                    const int32_t* __restrict__ invp0, int n,
                                                    int nrhs, const double* __restrict__ esc_row) {
  __shared__ T tile[64][65];
  const int o0 = blockIdx.x * 64, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int o = o0 + lane;
  const int rows = min(64, n - o0);
  const double s = (esc_row && o < n) ? esc_row[o] : 1.0;
  for (int j0 = 0; j0 < nrhs; j0 += 64) {
    const int jn = min(64, nrhs - j0);
    if (o < n) {
      for (int jj = wave; jj < jn; jj += 4) {
        double v = b[(int64_t)(j0 + jj) * n + o];
        if (esc_row) v *= s;
        tile[lane][jj] = (T)v;
      }
    }
    __syncthreads();
    for (int e = tid; e < rows * jn; e += 256) {
      const int r = e / jn, c = e - r * jn;
      Y[(int64_t)invp0[o0 + r] * nrhs + j0 + c] = tile[r][c];
    }
    __syncthreads();
  }
}

// X (row-major, permuted) -> x (column-major, original numbering), the
// transpose of k_gather_rhs's access pattern.  Grid: ceil(n / 64).
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_sol(const T* __restrict__ X,
                                                     double* __restrict__ x,
                                                     const int32_t* __restrict__ invp0, int n,
                                                     int nrhs, const double* __restrict__ esc_row,
                                                     int accumulate) {
  __shared__ T tile[64][65];
  const int o0 = blockIdx.x * 64, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int o = o0 + lane;
  const int rows = min(64, n - o0);
  const double s = (esc_row && o < n) ? esc_row[o] : 1.0;
  for (int j0 = 0; j0 < nrhs; j0 += 64) {
    const int jn = min(64, nrhs - j0);
    for (int e = tid; e < rows * jn; e += 256) {
      const int r = e / jn, c = e - r * jn;
      tile[r][c] = X[(int64_t)invp0[o0 + r] * nrhs + j0 + c];
    }
    __syncthreads();
    if (o < n) {
      for (int jj = wave; jj < jn; jj += 4) {
        double v = (double)tile[lane][jj];
        if (esc_row) v *= s;
        double* dst = x + (int64_t)(j0 + jj) * n + o;
        if (accumulate)
          *dst += v;
        else
          *dst = v;
      }
    }
    __syncthreads();
  }
}

}  // namespace dev
}  // namespace spk
