// Engine implementation: plan upload, factorization / solve launch sequences
// and FP64 iterative refinement around an FP32 factor.
#include "engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <stdexcept>

#include <mutex>

#include "kernels.hip"

namespace spk {

std::string hip_error_string(hipError_t e) { return std::string(hipGetErrorString(e)); }

#define HIP_OK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")");  \
  } while (0)

namespace {
template <typename T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HIP_OK(hipMalloc(&p, n * sizeof(T)));
  return (T*)p;
}
template <typename T>
T* dupload(const std::vector<T>& v, hipStream_t s) {
  T* p = dalloc<T>(v.size());
  if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return p;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}
inline unsigned grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
}
// k_gather_rhs / k_scatter_sol: one workgroup per 64 rows
inline unsigned grid_rows64(int n) { return (unsigned)std::max(1, (n + 63) / 64); }
enum TaskArr { TA_POTRF = 0, TA_TRSM, TA_INPANEL, TA_SCHUR, TA_SDIAG, TA_SOFF, TA_SOFFM, TA_FUSED, TA_GINV, TA_SCHUR_LPT, TA_WIDE, TA_WIDE_LPT };
}  // namespace

namespace {
// Per-column reductions of column-major n x nrhs vectors in two passes with
// a fixed summation order: kRedBlocks blocks per column write partial
// results, one block per column combines them.
__device__ __forceinline__ void block_reduce2(double& a, double& b, bool is_max) {
  __shared__ double sa[256], sb[256];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      if (is_max) {
        sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + s]);
        sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + s]);
      } else {
        sa[threadIdx.x] += sa[threadIdx.x + s];
      }
    }
    __syncthreads();
  }
  a = sa[0];
  b = sb[0];
  __syncthreads();
}

// partial max |r| and max |x| of column blockIdx.y over block blockIdx.x's rows
__global__ void k_colmax2_part(const double* __restrict__ r, const double* __restrict__ x, int n,
                               double* __restrict__ part) {
  const int j = blockIdx.y;
  const int64_t per = ((int64_t)n + kRedBlocks - 1) / kRedBlocks;
  const int64_t i0 = (int64_t)blockIdx.x * per, i1 = std::min<int64_t>(n, i0 + per);
  double mr = 0.0, mx = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    mr = fmax(mr, fabs(r[(int64_t)j * n + i]));
    mx = fmax(mx, fabs(x[(int64_t)j * n + i]));
  }
  block_reduce2(mr, mx, true);
  if (threadIdx.x == 0) {
    part[2 * ((int64_t)j * kRedBlocks + blockIdx.x)] = mr;
    part[2 * ((int64_t)j * kRedBlocks + blockIdx.x) + 1] = mx;
  }
}

// out[2j] = max |r_j|, out[2j+1] = max |x_j|
__global__ void k_colmax2_fin(const double* __restrict__ part, double* __restrict__ out) {
  const int j = blockIdx.x;
  const int t = threadIdx.x;
  double mr = t < kRedBlocks ? part[2 * ((int64_t)j * kRedBlocks + t)] : 0.0;
  double mx = t < kRedBlocks ? part[2 * ((int64_t)j * kRedBlocks + t) + 1] : 0.0;
  block_reduce2(mr, mx, true);
  if (t == 0) {
    out[2 * j] = mr;
    out[2 * j + 1] = mx;
  }
}

// partial sums of a_j . b_j
__global__ void k_coldot_part(const double* __restrict__ a, const double* __restrict__ b, int n,
                              double* __restrict__ part) {
  const int j = blockIdx.y;
  const int64_t per = ((int64_t)n + kRedBlocks - 1) / kRedBlocks;
  const int64_t i0 = (int64_t)blockIdx.x * per, i1 = std::min<int64_t>(n, i0 + per);
  double acc = 0.0, unused = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc += a[(int64_t)j * n + i] * b[(int64_t)j * n + i];
  block_reduce2(acc, unused, false);
  if (threadIdx.x == 0) part[(int64_t)j * kRedBlocks + blockIdx.x] = acc;
}

__global__ void k_coldot_fin(const double* __restrict__ part, double* __restrict__ out) {
  const int j = blockIdx.x;
  const int t = threadIdx.x;
  double acc = t < kRedBlocks ? part[(int64_t)j * kRedBlocks + t] : 0.0, unused = 0.0;
  block_reduce2(acc, unused, false);
  if (t == 0) out[j] = acc;
}

// CSR values from the caller's CSC values
__global__ void k_gather_vals(double* __restrict__ cval, const double* __restrict__ vals,
                              const int64_t* __restrict__ src, int64_t nnz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) cval[k] = vals[src[k]];
}

// out = b - A x (b given) or out = A x, A symmetric as full CSR, vectors
// column-major n x nrhs.  One thread per row, 16 right-hand sides per pass
// over the row's entries (each entry read once per 16 right-hand sides, 16
// independent x loads in flight): the lanes of a wave hold 64 consecutive
// rows, so for a banded (grid) matrix an entry position reads 64 nearly
// consecutive x values of each column.  The entries of a row are summed in
// CSR order; no atomics.
__global__ __launch_bounds__(256) void k_csr_mv(const int64_t* __restrict__ rowptr,
                                                const int32_t* __restrict__ col,
                                                const double* __restrict__ val,
                                                const double* __restrict__ x,
                                                const double* __restrict__ b,
                                                double* __restrict__ out, int n, int nrhs) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
  constexpr int JC = 16;
  for (int j0 = 0; j0 < nrhs; j0 += JC) {
    const int jn = min(JC, nrhs - j0);
    const double* __restrict__ xj = x + (int64_t)j0 * n;
    double acc[JC];
#pragma unroll
    for (int q = 0; q < JC; ++q) acc[q] = 0.0;
    // the next entry's value and column are loaded before this entry's x
    // values are used, so the column-index round trip overlaps them
    double v = p0 < p1 ? val[p0] : 0.0;
    int32_t c = p0 < p1 ? col[p0] : 0;
    for (int64_t p = p0; p < p1; ++p) {
      const bool more = p + 1 < p1;
      const double vn = more ? val[p + 1] : 0.0;
      const int32_t cn = more ? col[p + 1] : 0;
      const double* __restrict__ xc = xj + c;
#pragma unroll
      for (int q = 0; q < JC; ++q)
        if (q < jn) acc[q] += v * xc[(int64_t)q * n];
      v = vn;
      c = cn;
    }
#pragma unroll
    for (int q = 0; q < JC; ++q) {
      if (q < jn) {
        const int64_t o = (int64_t)(j0 + q) * n + row;
        out[o] = b ? b[o] - acc[q] : acc[q];
      }
    }
  }
}
// GMRES (refinement of the FP32 static LDL^T factor): column-major n x nrhs
// vectors, one Krylov space per right-hand side.
// part[((i nrhs + j) kRedBlocks + b)] = partial <V_i(:,j), w(:,j)> over block b's rows
__global__ void k_multidot_part(const double* __restrict__ V, int64_t vstride, const double* __restrict__ w,
                                int n, int nrhs, double* __restrict__ part) {
  const int j = blockIdx.y, i = blockIdx.z;
  const int64_t per = ((int64_t)n + kRedBlocks - 1) / kRedBlocks;
  const int64_t i0 = (int64_t)blockIdx.x * per, i1 = std::min<int64_t>(n, i0 + per);
  const double* __restrict__ v = V + (int64_t)i * vstride + (int64_t)j * n;
  const double* __restrict__ x = w + (int64_t)j * n;
  double acc = 0.0, unused = 0.0;
  for (int64_t r = i0 + threadIdx.x; r < i1; r += 256) acc += v[r] * x[r];
  block_reduce2(acc, unused, false);
  if (threadIdx.x == 0) part[((int64_t)i * nrhs + j) * kRedBlocks + blockIdx.x] = acc;
}

// out[i nrhs + j] = sum of the partials (fixed order)
__global__ void k_multidot_fin(const double* __restrict__ part, double* __restrict__ out) {
  const int64_t q = blockIdx.x;  // i nrhs + j
  const int t = threadIdx.x;
  double acc = t < kRedBlocks ? part[q * kRedBlocks + t] : 0.0, unused = 0.0;
  block_reduce2(acc, unused, false);
  if (t == 0) out[q] = acc;
}

// w(:,j) -= sum_{i < kv} h[i nrhs + j] V_i(:,j)
__global__ void k_gs_update(double* __restrict__ w, const double* __restrict__ V, int64_t vstride,
                            const double* __restrict__ h, int kv, int n, int nrhs) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n);
    double acc = 0.0;
    for (int i = 0; i < kv; ++i) acc += h[i * nrhs + j] * V[(int64_t)i * vstride + idx];
    w[idx] -= acc;
  }
}

// dst(:,j) = src(:,j) / sqrt(nrm2[j]) (0 where nrm2[j] <= 0)
__global__ void k_scale_cols(double* __restrict__ dst, const double* __restrict__ src,
                             const double* __restrict__ nrm2, int n, int nrhs) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const double q = nrm2[idx / n];
    dst[idx] = q > 0.0 ? src[idx] / sqrt(q) : 0.0;
  }
}

// x(:,j) += sum_{i < kv} y[i nrhs + j] Z_i(:,j)
__global__ void k_add_combo(double* __restrict__ x, const double* __restrict__ Z, int64_t zstride,
                            const double* __restrict__ y, int kv, int n, int nrhs) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n);
    double acc = 0.0;
    for (int i = 0; i < kv; ++i) acc += y[i * nrhs + j] * Z[(int64_t)i * zstride + idx];
    x[idx] += acc;
  }
}
}  // namespace

// per-RHS max |r| and max |x| into out[2j], out[2j+1]
void Engine::colmax2(const double* r, const double* x, int nrhs, double* out) {
  k_colmax2_part<<<dim3(kRedBlocks, (unsigned)nrhs), 256, 0, stream_>>>(r, x, n_, d_part_);
  k_colmax2_fin<<<(unsigned)nrhs, 256, 0, stream_>>>(d_part_, out);
}

// out[j] = a_j . b_j
void Engine::coldot(const double* a, const double* b, int nrhs, double* out) {
  k_coldot_part<<<dim3(kRedBlocks, (unsigned)nrhs), 256, 0, stream_>>>(a, b, n_, d_part_);
  k_coldot_fin<<<(unsigned)nrhs, 256, 0, stream_>>>(d_part_, out);
}

// y = A x on the full CSR (FP64, original numbering)
void Engine::csr_mv(const double* x, const double* b, double* out, int nrhs) {
  k_csr_mv<<<grid_for(n_), 256, 0, stream_>>>(d_crow_, d_ccol_, d_cval_, x, b, out, n_, nrhs);
}

Engine::Engine() { kind_stats.resize(13); }

Engine::~Engine() {
  try {
    free_device();
    destroy_graphs();
    for (auto& e : ev_pool_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  } catch (...) {
  }
}

void Engine::free_device() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  destroy_graphs();
  dfree(d_sn_);
  dfree(d_lindx0_);
  dfree(d_tcol_);
  dfree(d_tmap_);
  dfree(d_rel_);
  dfree(d_merges_);
  dfree(d_mmap_);
  dfree(d_amap_);
  dfree(d_invp0_);
  for (auto& t : d_tasks_) dfree(t);
  dfree(d_info_);
  dfree(d_fcnt_);
  dfree(d_df_fwd_);
  dfree(d_df_bwd_);
  dfree(d_df_fend_);
  dfree(d_df_bend_);
  dfree(d_df_tgt_);
  dfree(d_df_dcnt0_);
  dfree(d_df_rcnt0_);
  dfree(d_invb_);
  dfree(d_df_state_);
  dfree(d_acolptr_);
  dfree(d_crow_);
  dfree(d_ccol_);
  dfree(d_csrc_);
  dfree(d_cval_);
  dfree(d_arowind_);
  dfree(d_acolidx_);
  dfree(d_vals_);
  dfree(d_scale_);
  dfree(d_arena64_);
  dfree(d_arena32_);
  dfree(d_dinv_);
  dfree(d_garena_);
  ginv_ready_ = false;
  dfree(d_Y_);
  dfree(d_X_);
  dfree(d_Xf_);
  dfree(d_b_);
  dfree(d_x_);
  dfree(d_r_);
  dfree(d_red_);
  dfree(d_part_);
  dfree(d_z_);
  dfree(d_p_);
  dfree(d_q_);
  dfree(d_s_);
  dfree(d_gmV_);
  dfree(d_gmh_);
  gm_cap_ = gm_hcap_ = 0;
  cap_nrhs_ = 0;
}

void Engine::init(const EngineOptions& o) {
  // no device work here: the host analysis runs without a GPU
  opt_ = o;
}

int64_t Engine::device_bytes(int nrhs) const {
  auto vb = [](const auto& v) { return (int64_t)(v.size() * sizeof(v[0])); };
  const Plan& p = plan_;
  int64_t b = vb(p.sn) + vb(p.lindx0) + p.upd_size * 16 + p.rel_size * 4 + vb(p.merges) + vb(p.mmap) + vb(p.amap) + vb(p.perm0) +
              vb(p.potrf_tasks) + vb(p.trsm_tasks) + vb(p.inpanel_tasks) +
              2 * (p.n_schur + p.n_schur_wide) * (int64_t)sizeof(Task4) + vb(p.fused_tasks) +
              vb(p.sdiag_tasks) + vb(p.soffm_tasks);
  b += (int64_t)(n_ + 1) * 8 + nnzA_ * (4 + 4 + 8) + (int64_t)n_ * 8;  // CSC, values, scaling
  b += (int64_t)n_ * 4 + 4;                                            // invp0, pivot flag
  if (!p.fused_tasks.empty()) b += (int64_t)p.nsuper * 4;               // fused-factor load counters
  const int64_t elem = opt_.precision == 0 ? 8 : 4;
  b += p.arena_elems * elem + std::max<int64_t>(1, p.inv_tiles) * kSub * kSub * 8;
  if (p.ginv_ok && opt_.solve_inverse) b += p.arena_elems * elem + vb(p.gtiles);  // G and its tiles
  if (opt_.precision == 1)  // full CSR of A (row pointers, columns, sources, values), at most 2 nnz
    b += (int64_t)(n_ + 1) * 8 + 2 * nnzA_ * (4 + 8 + 8);
  b += 2 * (int64_t)std::max(1, nrhs) * kRedBlocks * 8;
  b += (int64_t)n_ * std::max(1, nrhs) * (3 * elem + (opt_.precision == 0 ? 3 : 6) * 8) +
       48 * std::max(1, nrhs);
  return b;
}

void Engine::release_device() {
  free_device();
  device_ready_ = have_values_ = factored_ = false;
  cap_nrhs_ = 0;
}

void Engine::reset_stats() {
  for (auto& k : kind_stats) k = KindStats();
  trace.clear();
}

void Engine::symbolic(Int n, const Ptr* colptr, const Idx* rowind, int baseval) {
  free_device();
  have_symbolic_ = have_values_ = factored_ = device_ready_ = false;
  const Int* up = opt_.user_perm.empty() ? nullptr : opt_.user_perm.data();
  symbolic_factorization(n, colptr, rowind, baseval, opt_.ordering, opt_.relax, up, sym_,
                         opt_.refine);
  build_plan(sym_, n, colptr, rowind, baseval, plan_, opt_.merge_updates);
  tuned_ = false;
  tune_log.clear();
  n_ = n;
  nnzA_ = colptr[n] - baseval;
  h_colptr_.assign(n + 1, 0);
  for (Int i = 0; i <= n; ++i) h_colptr_[i] = colptr[i] - baseval;
  h_rowind_.assign(nnzA_, 0);
  for (Int c = 0; c < n; ++c)
    for (int64_t p = h_colptr_[c]; p < h_colptr_[c + 1]; ++p) h_rowind_[p] = rowind[p] - baseval;
  have_symbolic_ = true;
}

// Device set-up on first numeric use: stream, plan, maps, task lists, arena.
void Engine::ensure_device() {
  HIP_OK(hipSetDevice(opt_.device));  // the calling thread may not have selected it
  if (!stream_) {
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  }
  if (device_ready_) return;
  hipStream_t s = stream_;
  std::vector<int32_t> colidx(nnzA_);
  for (Int c = 0; c < n_; ++c)
    for (int64_t p = h_colptr_[c]; p < h_colptr_[c + 1]; ++p) colidx[p] = c;
  d_sn_ = dupload(plan_.sn, s);
  d_lindx0_ = dupload(plan_.lindx0, s);
  {
    // scatter bases and relative indices of the Schur scatter, built on the
    // device
    RelGroup* groups = dupload(plan_.rel_groups, s);
    d_tcol_ = dalloc<int64_t>((size_t)std::max<int64_t>(1, plan_.upd_size));
    d_tmap_ = dalloc<int64_t>((size_t)std::max<int64_t>(1, plan_.upd_size));
    d_rel_ = dalloc<int32_t>((size_t)std::max<int64_t>(1, plan_.rel_size));
    int* bad = dalloc<int>(1);
    HIP_OK(hipMemsetAsync(bad, 0, sizeof(int), s));
    const int64_t ng = (int64_t)plan_.rel_groups.size();
    if (ng > 0) {
      dev::k_build_rel<<<(unsigned)std::min<int64_t>(ng, 65536), 256, 0, s>>>(
          (const SnInfo*)d_sn_, d_lindx0_, groups, ng, d_rel_, d_tcol_, d_tmap_, bad);
      HIP_OK(hipGetLastError());
    }
    int hbad = 0;
    HIP_OK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    dfree(groups);
    dfree(bad);
    if (hbad) throw std::logic_error("symbolic structure is not closed under elimination");
  }
  d_merges_ = dupload(plan_.merges, s);
  d_mmap_ = dupload(plan_.mmap, s);
  d_amap_ = dupload(plan_.amap, s);
  {
    // invp0[old] = new, 0-based: the right-hand side gather and the solution
    // scatter walk the original numbering
    std::vector<int32_t> invp0(plan_.perm0.size());
    for (size_t i = 0; i < invp0.size(); ++i) invp0[plan_.perm0[i]] = (int32_t)i;
    d_invp0_ = dupload(invp0, s);
  }
  d_tasks_[TA_POTRF] = dupload(plan_.potrf_tasks, s);
  d_tasks_[TA_TRSM] = dupload(plan_.trsm_tasks, s);
  d_tasks_[TA_INPANEL] = dupload(plan_.inpanel_tasks, s);
  {
    // Schur task lists, written on the device from one record per supernode
    for (int k : {TA_SCHUR, TA_SCHUR_LPT})
      d_tasks_[k] = dalloc<Task4>((size_t)std::max<int64_t>(1, plan_.n_schur));
    for (int k : {TA_WIDE, TA_WIDE_LPT})
      d_tasks_[k] = dalloc<Task4>((size_t)std::max<int64_t>(1, plan_.n_schur_wide));
    const int64_t ne = (int64_t)plan_.schur_emit.size();
    if (ne > 0) {
      SchurEmit* emit = dupload(plan_.schur_emit, s);
      dev::k_emit_schur<<<(unsigned)std::min<int64_t>(ne, 65536), 256, 0, s>>>(
          emit, ne, (Task4*)d_tasks_[TA_SCHUR], (Task4*)d_tasks_[TA_SCHUR_LPT], (Task4*)d_tasks_[TA_WIDE],
          (Task4*)d_tasks_[TA_WIDE_LPT]);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(s));
      dfree(emit);
    }
  }
  d_tasks_[TA_FUSED] = dupload(plan_.fused_tasks, s);
  d_tasks_[TA_SDIAG] = dupload(plan_.sdiag_tasks, s);
  d_tasks_[TA_SOFFM] = dupload(plan_.soffm_tasks, s);
  if (use_ginv()) d_tasks_[TA_GINV] = dupload(plan_.gtiles, s);
  d_info_ = dalloc<int>(1);
  if (!plan_.fused_tasks.empty()) {  // load counters of the fused block factor + TRSM
    d_fcnt_ = dalloc<int>((size_t)plan_.nsuper);
    // on the engine's stream: a legacy-stream memset would break another
    // thread's graph capture
    HIP_OK(hipMemsetAsync(d_fcnt_, 0, sizeof(int) * (size_t)plan_.nsuper, s));
  }
  if (plan_.df_ok && use_df()) {  // switched off: nothing is uploaded
    const DataflowLists df = build_dataflow_lists(plan_);
    df_nf_ = (int)df.fwd.size();
    df_nb_ = (int)df.bwd.size();
    d_df_fwd_ = dupload(df.fwd, s);
    d_df_bwd_ = dupload(df.bwd, s);
    d_df_fend_ = dupload(df.fwd_end, s);
    d_df_bend_ = dupload(df.bwd_end, s);
    d_df_tgt_ = dupload(df.tgt, s);
    d_df_dcnt0_ = dupload(df.dcnt0, s);
    d_df_rcnt0_ = dupload(df.rcnt0, s);
    d_invb_ = dupload(plan_.inv_base, s);
    d_df_state_ = dalloc<int>(3 + 4 * (size_t)plan_.nsuper);
    int cus = 0;
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, opt_.device));
    df_grid_ = std::max(1, 2 * cus);
  }
  d_acolptr_ = dupload(h_colptr_, s);
  d_arowind_ = dupload(h_rowind_, s);
  if (opt_.precision == 1) {
    // full CSR of the symmetric matrix: entry (r,c) of the lower CSC goes to
    // row r (column c) and, off the diagonal, to row c (column r)
    std::vector<int64_t> rp(n_ + 1, 0);
    for (Int c = 0; c < n_; ++c)
      for (int64_t p = h_colptr_[c]; p < h_colptr_[c + 1]; ++p) {
        const int32_t r = h_rowind_[p];
        ++rp[r + 1];
        if (r != c) ++rp[c + 1];
      }
    for (Int i = 0; i < n_; ++i) rp[i + 1] += rp[i];
    nnz_full_ = rp[n_];
    std::vector<int32_t> col(nnz_full_);
    std::vector<int64_t> src(nnz_full_);
    std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
    for (Int c = 0; c < n_; ++c)
      for (int64_t p = h_colptr_[c]; p < h_colptr_[c + 1]; ++p) {
        const int32_t r = h_rowind_[p];
        col[fill[r]] = c;
        src[fill[r]++] = p;
        if (r != c) {
          col[fill[c]] = r;
          src[fill[c]++] = p;
        }
      }
    d_crow_ = dupload(rp, s);
    d_ccol_ = dupload(col, s);
    d_csrc_ = dupload(src, s);
    d_cval_ = dalloc<double>((size_t)nnz_full_);
  }
  d_acolidx_ = dupload(colidx, s);
  d_vals_ = dalloc<double>(nnzA_);
  d_scale_ = dalloc<double>(n_);
  if (opt_.precision == 0)
    d_arena64_ = dalloc<double>(plan_.arena_elems);
  else
    d_arena32_ = dalloc<float>(plan_.arena_elems);
  // inverted 16x16 diagonal tiles (double-sized so both precisions fit)
  d_dinv_ = dalloc<double>(std::max<int64_t>(1, plan_.inv_tiles) * kSub * kSub);
  if (use_ginv()) {
    if (opt_.precision == 0)
      d_garena_ = dalloc<double>(plan_.arena_elems);
    else
      d_garena_ = dalloc<float>(plan_.arena_elems);
  }
  ginv_ready_ = false;
  HIP_OK(hipStreamSynchronize(s));
  device_ready_ = true;
}

void Engine::distribute(const double* nzval, bool on_device) {
  if (!have_symbolic_) throw std::logic_error("DistributeMatrix called before SymbolicFactorization");
  ensure_device();
  HIP_OK(hipMemcpyAsync(d_vals_, nzval, nnzA_ * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  if (d_cval_) {
    k_gather_vals<<<grid_for(nnz_full_), 256, 0, stream_>>>(d_cval_, d_vals_, d_csrc_, nnz_full_);
    HIP_OK(hipGetLastError());
  }
  // ||A||_inf of the symmetric matrix and the equilibration scaling
  // s_i = 1/sqrt(|a_ii|) used by the FP32 factor.
  std::vector<double> vals;
  const double* hv = nzval;
  if (on_device) {
    vals.resize(nnzA_);
    HIP_OK(hipMemcpyAsync(vals.data(), nzval, nnzA_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    hv = vals.data();
  }
  std::vector<double> rowsum(n_, 0.0), diag(n_, 0.0);
  for (Int c = 0; c < n_; ++c)
    for (int64_t p = h_colptr_[c]; p < h_colptr_[c + 1]; ++p) {
      int32_t r = h_rowind_[p];
      double a = std::fabs(hv[p]);
      rowsum[r] += a;
      if (r != c)
        rowsum[c] += a;
      else
        diag[c] = hv[p];
    }
  normA_inf_ = 0.0;
  for (double v : rowsum) normA_inf_ = std::max(normA_inf_, v);
  std::vector<double> sc(n_, 1.0);
  if (opt_.precision == 1)
    for (Int i = 0; i < n_; ++i) sc[i] = diag[i] != 0.0 ? 1.0 / std::sqrt(std::fabs(diag[i])) : 1.0;
  HIP_OK(hipMemcpyAsync(d_scale_, sc.data(), n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
  have_values_ = true;
  factored_ = false;
  fp32_level_ = 0;
}

// Equilibrated assembly: entry (r,c) scaled by s_r s_c, plus `shift` on the
// diagonal (only after an FP32 breakdown of an SPD matrix).
template <typename T>
__global__ void k_assemble_scaled(T* __restrict__ arena, const int64_t* __restrict__ amap,
                                  const double* __restrict__ vals, const int32_t* __restrict__ rowind,
                                  const int32_t* __restrict__ colidx, const double* __restrict__ s,
                                  int64_t nnz, double shift) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t dst = amap[e];
    if (dst < 0) continue;
    const int r = rowind[e], c = colidx[e];
    double v = vals[e] * s[r] * s[c];
    if (r == c) v += shift;
    arena[dst] = (T)v;
  }
}

template <typename T>
void Engine::assemble(T* arena) {
  dev::k_fill_zero<T><<<grid_for(plan_.arena_elems), 256, 0, stream_>>>(arena, plan_.arena_elems);
  if (opt_.precision == 0)
    dev::k_assemble<T><<<grid_for(nnzA_), 256, 0, stream_>>>(arena, d_amap_, d_vals_, nullptr, nnzA_);
  else
    k_assemble_scaled<T><<<grid_for(nnzA_), 256, 0, stream_>>>(arena, d_amap_, d_vals_, d_arowind_,
                                                               d_acolidx_, d_scale_, nnzA_, fp32_shift_);
  HIP_OK(hipGetLastError());
}

void Engine::launch_begin(int kind, int64_t grid, hipStream_t st, int level) {
  if (!st) st = stream_;
  if (capturing_)
    cap_kinds_.push_back(kind);
  else
    kind_stats[kind].launches++;
  if (!opt_.profile || !(profile_mask_ & (1u << kind))) return;
  ev_grid_.push_back(grid);
  ev_level_.push_back(level);
  while (ev_pool_.size() < ev_next_ + 2) {
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    ev_pool_.push_back(e);
  }
  hipEvent_t a = ev_pool_[ev_next_], b = ev_pool_[ev_next_ + 1];
  ev_next_ += 2;
  HIP_OK(hipEventRecord(a, st));
  ev_open_.push_back({kind, {a, b}});
}

void Engine::launch_end(int kind, hipStream_t st) {
  if (!st) st = stream_;
  if (!opt_.profile || !(profile_mask_ & (1u << kind))) return;
  HIP_OK(hipEventRecord(ev_open_.back().second.second, st));
}

void Engine::destroy_graph(Graph& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (auto& e : g.events) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  g = Graph();
}

void Engine::destroy_graphs() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  pending_graphs_.clear();
  destroy_graph(fgraph_);
  for (auto& g : sgraphs_) destroy_graph(g);
  sgraphs_.clear();
}

// Runs `body` (a sequence of launches on stream_) through the graph slot g:
// captured on first use or when the key changes, replayed afterwards.
// Profile mode launches eagerly: events recorded inside a graph cannot be
// timed with hipEventElapsedTime.
template <typename F>
void Engine::run_graph(Graph& g, int64_t key, F&& body) {
  if (!use_graphs_ || opt_.profile) {
    body();
    return;
  }
  if (!g.exec || g.key != key) {
    destroy_graph(g);
    cap_kinds_.clear();
    cap_events_.clear();
    cap_grids_.clear();
    hipGraph_t graph = nullptr;
    HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    capturing_ = true;
    try {
      body();
    } catch (...) {
      capturing_ = false;
      (void)hipStreamEndCapture(stream_, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      for (auto& e : cap_events_) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
      }
      cap_events_.clear();
      throw;
    }
    capturing_ = false;
    HIP_OK(hipStreamEndCapture(stream_, &graph));
    hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    g.events.swap(cap_events_);
    g.grids.swap(cap_grids_);
    g.kinds.swap(cap_kinds_);
    if (e != hipSuccess) {
      g.exec = nullptr;
      destroy_graph(g);
      HIP_OK(e);
    }
    g.key = key;
  }
  g.used = ++graph_clock_;
  HIP_OK(hipGraphLaunch(g.exec, stream_));
  for (int k : g.kinds) kind_stats[k].launches++;
  if (!g.events.empty()) pending_graphs_.push_back(&g);
}

Engine::Graph& Engine::solve_graph(int64_t key) {
  for (auto& g : sgraphs_)
    if (g.key == key) return g;
  constexpr size_t kMaxSolveGraphs = 4;
  sgraphs_.reserve(kMaxSolveGraphs);  // pending_graphs_ keeps pointers into it
  if (sgraphs_.size() < kMaxSolveGraphs) {
    sgraphs_.emplace_back();
    return sgraphs_.back();
  }
  Graph* lru = &sgraphs_[0];
  for (auto& g : sgraphs_)
    if (g.used < lru->used) lru = &g;
  if (stream_) HIP_OK(hipStreamSynchronize(stream_));
  pending_graphs_.clear();
  destroy_graph(*lru);
  return *lru;
}

void Engine::collect_events() {
  if (!opt_.profile) return;
  for (Graph* g : pending_graphs_)
    for (size_t i = 0; i < g->events.size(); ++i) {
      auto& o = g->events[i];
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, o.second.first, o.second.second));
      kind_stats[o.first].ms += ms;
      if (trace.size() < 100000) trace.push_back({o.first, g->grids[i], ms, -1});
    }
  pending_graphs_.clear();
  for (size_t i = 0; i < ev_open_.size(); ++i) {
    auto& o = ev_open_[i];
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, o.second.first, o.second.second));
    kind_stats[o.first].ms += ms;
    if (trace.size() < 100000)
      trace.push_back({o.first, i < ev_grid_.size() ? ev_grid_[i] : 0, ms, i < ev_level_.size() ? ev_level_[i] : -1});
  }
  ev_open_.clear();
  ev_grid_.clear();
  ev_level_.clear();
  ev_next_ = 0;
}

template <typename T, bool LDL, int NBM>
static void diag_launch(unsigned g, hipStream_t st, const SnInfo* sn, const Task4* t, T* arena, T* dinv,
                        int* info) {
  dev::k_diag<T, LDL, NBM><<<g, 256, 0, st>>>(sn, t, arena, dinv, info);
}
template <typename T, bool LDL, int NBM>
static void ptrsm_launch(unsigned g, hipStream_t st, const SnInfo* sn, const Task4* t, T* arena,
                         const T* dinv) {
  dev::k_ptrsm<T, LDL, NBM><<<g, 512, 0, st>>>(sn, t, arena, dinv);
}

template <typename T>
void Engine::launch_diag(int bucket, bool ldl, unsigned g, const Task4* t, T* arena) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  T* dinv = (T*)d_dinv_;
#define SPK_DIAG(NBM)                                                 \
  case NBM:                                                           \
    if (ldl)                                                          \
      diag_launch<T, true, NBM>(g, stream_, sn, t, arena, dinv, d_info_);  \
    else                                                              \
      diag_launch<T, false, NBM>(g, stream_, sn, t, arena, dinv, d_info_); \
    break;
  switch (bucket) {
    SPK_DIAG(1)
    SPK_DIAG(2)
    SPK_DIAG(4)
    SPK_DIAG(6)
    SPK_DIAG(10)
    default:
      throw std::logic_error("bad diagonal-block bucket");
  }
#undef SPK_DIAG
}

template <typename T>
void Engine::launch_ptrsm(int bucket, bool ldl, unsigned g, const Task4* t, T* arena) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const T* dinv = (const T*)d_dinv_;
#define SPK_PTRSM(NBM)                                            \
  case NBM:                                                       \
    if (ldl)                                                      \
      ptrsm_launch<T, true, NBM>(g, stream_, sn, t, arena, dinv);  \
    else                                                          \
      ptrsm_launch<T, false, NBM>(g, stream_, sn, t, arena, dinv); \
    break;
  switch (bucket) {
    SPK_PTRSM(1)
    SPK_PTRSM(2)
    SPK_PTRSM(4)
    SPK_PTRSM(6)
    SPK_PTRSM(10)
    default:
      throw std::logic_error("bad diagonal-block bucket");
  }
#undef SPK_PTRSM
}

template <typename T, bool LDL, int NBM>
static void diagtrsm_launch(unsigned g, hipStream_t st, const SnInfo* sn, const Task4* t, T* arena,
                            T* dinv, int* info, int* cnt) {
  dev::k_diagtrsm<T, LDL, NBM><<<g, 256, 0, st>>>(sn, t, arena, dinv, info, cnt);
}

template <typename T>
void Engine::launch_diagtrsm(int bucket, bool ldl, unsigned g, const Task4* t, T* arena) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  T* dinv = (T*)d_dinv_;
#define SPK_FUSED(NBM)                                                                   \
  case NBM:                                                                              \
    if (ldl)                                                                             \
      diagtrsm_launch<T, true, NBM>(g, stream_, sn, t, arena, dinv, d_info_, d_fcnt_);  \
    else                                                                                 \
      diagtrsm_launch<T, false, NBM>(g, stream_, sn, t, arena, dinv, d_info_, d_fcnt_); \
    break;
  switch (bucket) {
    SPK_FUSED(1)
    SPK_FUSED(2)
    SPK_FUSED(4)
    SPK_FUSED(6)
    SPK_FUSED(10)
    default:
      throw std::logic_error("bad diagonal-block bucket");
  }
#undef SPK_FUSED
}

// One Schur launch of range r: 64 x 64 tiles staged through LDS (k_update),
// 128 x 64 tiles (k_schur_wide, alt bit 1) or 64 x 64 tiles with one wave per
// tile (k_schur_w1, alt bit 2), plan or longest-first task order (alt bit 0),
// each tile's k chunks split over ks workgroups (waves for k_schur_w1).
template <typename T>
void Engine::launch_schur(const LaunchRange& r, int alt, int ks, T* arena, hipStream_t st) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const bool ldl = opt_.ldl != 0;
  const bool wide = (alt & 2) != 0;
  const int64_t cnt = wide ? r.count_w : r.count;
  if (cnt == 0) return;
  const Task4* t = (const Task4*)d_tasks_[wide ? ((alt & 1) ? TA_WIDE_LPT : TA_WIDE)
                                              : ((alt & 1) ? TA_SCHUR_LPT : TA_SCHUR)] +
                   (wide ? r.first_w : r.first);
  const unsigned gs = (unsigned)(cnt * ks);
  // k_schur_w1 addresses panels through 32-bit buffer offsets: below 2 GiB
  if ((alt & 4) && plan_.max_panel * (int64_t)sizeof(T) >= (int64_t(1) << 31)) alt &= 1;
  if (alt & 4) {  // one wave per (task, part): four per workgroup
    const unsigned gw = (unsigned)((cnt * ks + 3) / 4);
    if (ldl)
      dev::k_schur_w1<T, true><<<gw, 256, 0, st>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, cnt, ks);
    else
      dev::k_schur_w1<T, false><<<gw, 256, 0, st>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, cnt, ks);
  } else if (wide) {
    if (ldl)
      dev::k_schur_wide<T, true><<<gs, 256, 0, st>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, cnt, ks);
    else
      dev::k_schur_wide<T, false><<<gs, 256, 0, st>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, cnt, ks);
  } else {
    if (ldl)
      dev::k_update<T, 1, true><<<gs, 256, 0, st>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, cnt, ks);
    else
      dev::k_update<T, 1, false><<<gs, 256, 0, st>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, cnt, ks);
  }
}

template <typename T>
void Engine::factor_sequence(T* arena) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const bool ldl = opt_.ldl != 0;
  for (const LaunchRange& r : plan_.factor_seq) {
    const unsigned g = (unsigned)r.count;
    hipStream_t st = stream_;
    launch_begin(r.kind, r.count, st, r.level);
    switch (r.kind) {
      case 0: {
        const Task4* t = (const Task4*)d_tasks_[TA_POTRF] + r.first;
        launch_diag<T>(r.aux, ldl, g, t, arena);
        break;
      }
      case 7: {
        const Task4* t = (const Task4*)d_tasks_[TA_TRSM] + r.first;
        launch_ptrsm<T>(r.aux, ldl, g, t, arena);
        break;
      }
      case 9: {
        const Task4* t = (const Task4*)d_tasks_[TA_FUSED] + r.first;
        launch_diagtrsm<T>(r.aux, ldl, g, t, arena);
        break;
      }
      case 1: {
        const Task4* t = (const Task4*)d_tasks_[TA_INPANEL] + r.first;
        if (ldl)
          dev::k_update<T, 0, true><<<g, 256, 0, stream_>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, r.count, 1);
        else
          dev::k_update<T, 0, false><<<g, 256, 0, stream_>>>(sn, t, arena, d_tcol_, d_tmap_, d_rel_, d_merges_, d_mmap_, r.count, 1);
        break;
      }
      case 2:
        launch_schur<T>(r, r.alt, std::max(1, r.aux), arena, st);
        break;
      default:
        throw std::logic_error("bad factor launch kind");
    }
    launch_end(r.kind, st);
  }
  HIP_OK(hipGetLastError());
}

// Autotuner of the Schur launches (one per etree level).  Candidates: the
// plan's task order or longest-first, times a k-chunk split of 1, 2, 3, 4, 6
// or 8 workgroups per tile (at most the fewest chunks of any tile of the
// range, and at most 8192 workgroups).  Each candidate is launched three
// times on the arena, timed with HIP events on the library's stream, and the
// fastest minimum wins; the default configuration keeps its place unless
// another is at least 3% faster.  The Schur kernel only reads panels and
// adds into the arena, so the launches leave garbage that the
// factorization's assembly (fill + scatter of A) overwrites.
template <typename T>
void Engine::autotune_schur(T* arena) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const bool ldl = opt_.ldl != 0;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};
  auto launch = [&](const LaunchRange& r, int alt, int ks) { launch_schur<T>(r, alt, ks, arena, stream_); };
  auto time_of = [&](const LaunchRange& r, int alt, int ks) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      HIP_OK(hipEventRecord(e0, stream_));
      launch(r, alt, ks);
      HIP_OK(hipEventRecord(e1, stream_));
      HIP_OK(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    HIP_OK(hipGetLastError());
    return (double)best;
  };
  tune_log.clear();
  bool warm = false;
  for (LaunchRange& r : plan_.factor_seq) {
    if (r.kind != 2 || r.count == 0) continue;
    const int nqmin = r.nqmin;
    if (!warm) {  // first launches of the kernels in this process: load them untimed
      launch(r, 0, std::max(1, r.aux));
      if (r.count_w > 0) launch(r, 2, 1);
      launch(r, 4, 1);
      warm = true;
    }
    TuneRecord rec;
    rec.level = r.level;
    rec.tiles = r.count;
    rec.split0 = rec.split = std::max(1, r.aux);
    rec.alt0 = rec.alt = r.alt;
    rec.ms0 = rec.ms = time_of(r, r.alt, rec.split0);
    rec.candidates = 1;
    for (int alt = 0; alt < 6; ++alt)
      for (int ks : {1, 2, 3, 4, 6, 8}) {
        const int64_t tiles = (alt & 2) ? r.count_w : r.count;
        if (tiles == 0 || ks > nqmin || tiles * ks > 8192) continue;
        if (alt == rec.alt0 && ks == rec.split0) continue;
        const double ms = time_of(r, alt, ks);
        ++rec.candidates;
        if (ms < rec.ms && ms < 0.97 * rec.ms0) {
          rec.ms = ms;
          rec.split = ks;
          rec.alt = alt;
        }
      }
    r.aux = rec.split;
    r.alt = rec.alt;
    tune_log.push_back(rec);
  }
  HIP_OK(hipStreamSynchronize(stream_));
}

int Engine::factorize() {
  if (!have_values_) throw std::logic_error("Factorize called before DistributeMatrix");
  HIP_OK(hipSetDevice(opt_.device));
  auto t0 = std::chrono::steady_clock::now();
  const int big = INT_MAX;
  HIP_OK(hipMemcpyAsync(d_info_, &big, sizeof(int), hipMemcpyHostToDevice, stream_));
  ginv_ready_ = false;
  if (opt_.schur_kernel != 0 && !tuned_) {
    // one Schur kernel for every launch, plan order and k split
    for (LaunchRange& r : plan_.factor_seq)
      if (r.kind == 2)
        r.alt = opt_.schur_kernel == 3 ? 4 : (opt_.schur_kernel == 2 && r.count_w > 0) ? 2 : 0;
    tuned_ = true;
    destroy_graphs();
  }
  if (opt_.autotune && !tuned_ && plan_.schur_flops >= kAutotuneMinFlops) {
    if (opt_.precision == 0)
      autotune_schur<double>(d_arena64_);
    else
      autotune_schur<float>(d_arena32_);
    tuned_ = true;
    destroy_graphs();  // the captured factor sequence holds the old configuration
  }
  if (opt_.precision == 0) {
    run_graph(fgraph_, 1, [&] {
      assemble(d_arena64_);
      factor_sequence(d_arena64_);
    });
  } else {
    // FP32 factor of the equilibrated matrix.  An SPD matrix whose FP32
    // Cholesky breaks down is factored again with the diagonal of S A S
    // shifted by 1e-6, 1e-4, 1e-2: the shifted factor is still a
    // preconditioner for the FP64 refinement, which solves the unshifted
    // system.
    for (int level = fp32_level_;; ++level) {
      fp32_level_ = level;
      fp32_shift_ = level == 0 ? 0.0 : std::pow(10.0, -8.0 + 2.0 * level);
      run_graph(fgraph_, 2 + 16 * level, [&] {
        assemble(d_arena32_);
        factor_sequence(d_arena32_);
      });
      if (opt_.ldl || level == 3) break;
      int inf = 0;
      HIP_OK(hipMemcpyAsync(&inf, d_info_, sizeof(int), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      if (inf == big) break;
      HIP_OK(hipMemcpyAsync(d_info_, &big, sizeof(int), hipMemcpyHostToDevice, stream_));
    }
  }
  int info = 0;
  HIP_OK(hipMemcpyAsync(&info, d_info_, sizeof(int), hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
  collect_events();
  last_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  factored_ = (info == big);
  return factored_ ? 0 : info;
}

void Engine::ensure_solve_capacity(int nrhs) {
  if (nrhs <= cap_nrhs_) return;
  // captured solves refer to the old buffers
  for (auto& g : sgraphs_) destroy_graph(g);
  sgraphs_.clear();
  pending_graphs_.clear();
  dfree(d_Y_);
  dfree(d_X_);
  dfree(d_Xf_);
  dfree(d_b_);
  dfree(d_x_);
  dfree(d_r_);
  dfree(d_red_);
  dfree(d_part_);
  dfree(d_z_);
  dfree(d_p_);
  dfree(d_q_);
  dfree(d_s_);
  size_t tsz = opt_.precision == 0 ? sizeof(double) : sizeof(float);
  size_t cnt = (size_t)n_ * nrhs;
  HIP_OK(hipMalloc(&d_Y_, std::max<size_t>(1, cnt) * tsz));
  HIP_OK(hipMalloc(&d_X_, std::max<size_t>(1, cnt) * tsz));
  if (use_ginv()) HIP_OK(hipMalloc(&d_Xf_, std::max<size_t>(1, cnt) * tsz));
  d_b_ = dalloc<double>(cnt);
  d_x_ = dalloc<double>(cnt);
  d_r_ = dalloc<double>(cnt);
  d_red_ = dalloc<double>(2 * (size_t)nrhs);
  d_part_ = dalloc<double>(2 * (size_t)nrhs * kRedBlocks);
  if (opt_.precision == 1) {
    d_z_ = dalloc<double>(cnt);
    d_p_ = dalloc<double>(cnt);
    d_q_ = dalloc<double>(cnt);
    d_s_ = dalloc<double>(4 * (size_t)nrhs);
  }
  cap_nrhs_ = nrhs;
}

template <typename T, bool BWD>
void Engine::launch_sdiag(int bucket, unsigned g, const Task4* t, const T* arena, const T* Y, T* X, int nrhs) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const T* dinv = (const T*)d_dinv_;
  if (nrhs <= 16) {
    // one right-hand-side group: blocks spread over the waves
    switch (bucket) {
      case 1: dev::k_sdiag4<T, BWD, 1><<<g, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
      case 2: dev::k_sdiag4<T, BWD, 2><<<g, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
      case 4: dev::k_sdiag4<T, BWD, 4><<<g, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
      case 6: dev::k_sdiag4<T, BWD, 6><<<g, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
      case 10: dev::k_sdiag4<T, BWD, 10><<<g, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
      default: throw std::logic_error("bad diagonal-block bucket");
    }
    return;
  }
  const dim3 grid(g, (unsigned)((nrhs + 63) / 64));
  switch (bucket) {
    case 1: dev::k_sdiag3<T, BWD, 1><<<grid, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
    case 2: dev::k_sdiag3<T, BWD, 2><<<grid, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
    case 4: dev::k_sdiag3<T, BWD, 4><<<grid, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
    case 6: dev::k_sdiag3<T, BWD, 6><<<grid, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
    case 10: dev::k_sdiag3<T, BWD, 10><<<grid, 256, 0, stream_>>>(sn, t, arena, dinv, Y, X, nrhs); break;
    default: throw std::logic_error("bad diagonal-block bucket");
  }
}

// Dependency-driven sweeps: reset the counters, one persistent forward
// launch, (LDL) the D^{-1} scaling, one persistent backward launch.
template <typename T>
void Engine::solve_sequence_df(const T* arena, T* Y, T* X, int nrhs) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const int ns = plan_.nsuper;
  dev::DfState st;
  st.head = d_df_state_;
  st.dcnt = d_df_state_ + 2;
  st.dflag = st.dcnt + ns;
  st.rcnt = st.dflag + ns;
  st.bflag = st.rcnt + ns;
  st.abort_ = st.bflag + ns;
  dev::k_df_reset<<<grid_for(ns), 256, 0, stream_>>>(st, d_df_dcnt0_, d_df_rcnt0_, ns);
  const int nf = df_nf_, nb = df_nb_;
  launch_begin(3, nf);
  dev::k_fsolve_df<T><<<(unsigned)std::min(df_grid_, nf), 256, 0, stream_>>>(
      sn, d_df_fwd_, d_df_fend_, nf, d_df_tgt_, arena, (const T*)d_dinv_, d_lindx0_, d_invb_, Y, X,
      nrhs, st);
  launch_end(3);
  if (opt_.ldl) {
    launch_begin(3);
    dev::k_scale_d<T><<<(unsigned)ns, 256, 0, stream_>>>(sn, (const Task4*)d_tasks_[TA_SDIAG], arena, X, nrhs);
    launch_end(3);
  }
  launch_begin(6, nb);
  dev::k_bsolve_df<T><<<(unsigned)std::min(df_grid_, nb), 256, 0, stream_>>>(
      sn, d_df_bwd_, d_df_bend_, nb, d_df_tgt_, arena, (const T*)d_dinv_, d_lindx0_, d_invb_, X, nrhs,
      st);
  launch_end(6);
  HIP_OK(hipGetLastError());
}

template <typename T>
void Engine::ensure_ginv(const T* arena) {
  if (!use_ginv() || ginv_ready_) return;
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const Task4* tiles = (const Task4*)d_tasks_[TA_GINV];
  T* G = (T*)d_garena_;
  const T* dinv = (const T*)d_dinv_;
  for (const LaunchRange& r : plan_.gprep_seq) {
    const unsigned g = (unsigned)r.count;
    launch_begin(12, r.count, nullptr, -1);
    switch (r.aux) {
      case 1: dev::k_gprep<T, 1><<<g, 256, 0, stream_>>>(sn, tiles + r.first, arena, dinv, G); break;
      case 2: dev::k_gprep<T, 2><<<g, 256, 0, stream_>>>(sn, tiles + r.first, arena, dinv, G); break;
      case 4: dev::k_gprep<T, 4><<<g, 256, 0, stream_>>>(sn, tiles + r.first, arena, dinv, G); break;
      case 6: dev::k_gprep<T, 6><<<g, 256, 0, stream_>>>(sn, tiles + r.first, arena, dinv, G); break;
      case 10: dev::k_gprep<T, 10><<<g, 256, 0, stream_>>>(sn, tiles + r.first, arena, dinv, G); break;
      default: throw std::logic_error("bad diagonal-block bucket");
    }
    launch_end(12);
  }
  HIP_OK(hipGetLastError());
  ginv_ready_ = true;
}

// Partitioned-inverse sweeps: X zeroed, one forward launch per level
// (Y accumulates the updates, Xf receives X_I), one backward launch per
// level (partial sums into X).
template <typename T>
void Engine::ginv_sequence(T* Y, T* X, T* Xf, int nrhs) {
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const Task4* tiles = (const Task4*)d_tasks_[TA_GINV];
  const T* G = (const T*)d_garena_;
  const T* arena = opt_.precision == 0 ? (const T*)(void*)d_arena64_ : (const T*)(void*)d_arena32_;
  const bool ldl = opt_.ldl != 0;
  const int64_t cnt = (int64_t)n_ * nrhs;
  dev::k_fill_zero<T><<<grid_for(cnt), 256, 0, stream_>>>(X, cnt);
  // right-hand-side blocks of 16 per workgroup: a launch with at least
  // kGinvGrid tiles walks every block in each workgroup (G read once); one
  // with fewer spreads the blocks over up to kGinvGrid workgroups
  const int64_t nblk = (nrhs + 15) / 16;
  auto grid_y = [&](int64_t tiles) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk, (kGinvGrid + tiles - 1) / tiles));
  };
  for (const LaunchRange& r : plan_.gfwd_seq) {
    const dim3 grid((unsigned)r.count, grid_y(r.count));
    const Task4* t = tiles + r.first;
    launch_begin(10, r.count, nullptr, r.level);
#define SPK_GF(NBM)                                                                                        \
  case NBM:                                                                                                \
    if (ldl)                                                                                               \
      dev::k_gsolve_fwd<T, true, NBM><<<grid, 256, 0, stream_>>>(sn, t, G, arena, d_lindx0_, Y, Xf, nrhs); \
    else                                                                                                   \
      dev::k_gsolve_fwd<T, false, NBM><<<grid, 256, 0, stream_>>>(sn, t, G, arena, d_lindx0_, Y, Xf, nrhs); \
    break;
    switch (r.aux) {
      SPK_GF(1)
      SPK_GF(2)
      SPK_GF(4)
      SPK_GF(6)
      SPK_GF(10)
      default:
        throw std::logic_error("bad diagonal-block bucket");
    }
#undef SPK_GF
    launch_end(10);
  }
  for (const LaunchRange& r : plan_.gbwd_seq) {
    const dim3 grid((unsigned)r.count, grid_y(r.count));
    const Task4* t = tiles + r.first;
    launch_begin(11, r.count, nullptr, r.level);
    switch (r.aux) {
      case 1: dev::k_gsolve_bwd<T, 1><<<grid, 256, 0, stream_>>>(sn, t, G, d_lindx0_, Xf, X, nrhs); break;
      case 2: dev::k_gsolve_bwd<T, 2><<<grid, 256, 0, stream_>>>(sn, t, G, d_lindx0_, Xf, X, nrhs); break;
      case 4: dev::k_gsolve_bwd<T, 4><<<grid, 256, 0, stream_>>>(sn, t, G, d_lindx0_, Xf, X, nrhs); break;
      case 6: dev::k_gsolve_bwd<T, 6><<<grid, 256, 0, stream_>>>(sn, t, G, d_lindx0_, Xf, X, nrhs); break;
      case 10: dev::k_gsolve_bwd<T, 10><<<grid, 256, 0, stream_>>>(sn, t, G, d_lindx0_, Xf, X, nrhs); break;
      default: throw std::logic_error("bad diagonal-block bucket");
    }
    launch_end(11);
  }
  HIP_OK(hipGetLastError());
}

template <typename T>
void Engine::solve_sequence(const T* arena, T* Y, T* X, int nrhs, T* Xf) {
  if (use_ginv()) {
    if (!ginv_ready_) throw std::logic_error("partitioned-inverse solve used before its preparation");
    ginv_sequence<T>(Y, X, Xf ? Xf : (T*)d_Xf_, nrhs);
    return;
  }
  if (use_df()) {
    solve_sequence_df<T>(arena, Y, X, nrhs);
    return;
  }
  const SnInfo* sn = (const SnInfo*)d_sn_;
  const bool ldl = opt_.ldl != 0;
  const Task4* sd = (const Task4*)d_tasks_[TA_SDIAG];
  const Task4* sm = (const Task4*)d_tasks_[TA_SOFFM];
  for (const LaunchRange& r : plan_.fwd_seq) {
    const unsigned g = (unsigned)r.count;
    launch_begin(r.kind, r.count, nullptr, r.level);
    if (r.kind == 3) {
      if (r.aux > 0)
        launch_sdiag<T, false>(r.aux, g, sd + r.first, arena, Y, X, nrhs);
      else if (ldl)
        dev::k_sdiag_fwd<T, true><<<g, 256, 0, stream_>>>(sn, sd + r.first, arena, Y, X, nrhs);
      else
        dev::k_sdiag_fwd<T, false><<<g, 256, 0, stream_>>>(sn, sd + r.first, arena, Y, X, nrhs);
    } else {
      const dim3 grid((unsigned)r.count, (unsigned)((nrhs + 15) / 16));
      dev::k_soffm_fwd<T><<<grid, 256, 0, stream_>>>(sn, sm + r.first, arena, d_lindx0_, X, Y, nrhs);
    }
    launch_end(r.kind);
  }
  if (ldl) {
    launch_begin(3);
    dev::k_scale_d<T><<<(unsigned)plan_.nsuper, 256, 0, stream_>>>(sn, sd, arena, X, nrhs);
    launch_end(3);
  }
  for (const LaunchRange& r : plan_.bwd_seq) {
    const unsigned g = (unsigned)r.count;
    launch_begin(r.kind, r.count, nullptr, r.level);
    if (r.kind == 5) {
      const dim3 grid((unsigned)r.count, (unsigned)((nrhs + 15) / 16));
      dev::k_soffm_bwd<T><<<grid, 256, 0, stream_>>>(sn, sm + r.first, arena, d_lindx0_, X, nrhs);
    } else {
      if (r.aux > 0)
        launch_sdiag<T, true>(r.aux, g, sd + r.first, arena, X, X, nrhs);
      else if (ldl)
        dev::k_sdiag_bwd<T, true><<<g, 256, 0, stream_>>>(sn, sd + r.first, arena, X, nrhs);
      else
        dev::k_sdiag_bwd<T, false><<<g, 256, 0, stream_>>>(sn, sd + r.first, arena, X, nrhs);
    }
    launch_end(r.kind);
  }
  HIP_OK(hipGetLastError());
}

// r = b - A x (FP64, original numbering; FP32 factor mode, which holds A as
// full CSR), the residual check_solution forms (driver/utils.hpp:95-142).
void Engine::residual(const double* b, const double* x, double* r, int nrhs) {
  if (!d_cval_) throw std::logic_error("residual: the full CSR of A exists in FP32 factor mode only");
  csr_mv(x, b, r, nrhs);
  HIP_OK(hipGetLastError());
}



// CG step: alpha_j = rho_j / (p_j . q_j); x += alpha p; r -= alpha q.
__global__ void k_pcg_xr(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                         const double* __restrict__ q, const double* __restrict__ rho,
                         const double* __restrict__ pq, int n, int nrhs) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n);
    const double alpha = pq[j] != 0.0 ? rho[j] / pq[j] : 0.0;
    x[idx] += alpha * p[idx];
    r[idx] -= alpha * q[idx];
  }
}

// CG step: beta_j = rho_new_j / rho_j; p = z + beta p.
__global__ void k_pcg_p(double* __restrict__ p, const double* __restrict__ z,
                        const double* __restrict__ rho_new, const double* __restrict__ rho, int n,
                        int nrhs) {
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)n * nrhs;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n);
    const double beta = rho[j] != 0.0 ? rho_new[j] / rho[j] : 0.0;
    p[idx] = z[idx] + beta * p[idx];
  }
}

// One refinement step: d = S (SAS)^{-1} S r in the factor's precision,
// x += d, r = b - A x and the per-RHS norms, all on the device.
template <typename T>
void Engine::ir_step(T* Y, T* X, int nrhs) {
  dev::k_gather_rhs<T><<<grid_rows64(n_), 256, 0, stream_>>>(d_r_, Y, d_invp0_, n_, nrhs, d_scale_);
  solve_sequence<T>(d_arena32_, Y, X, nrhs);
  dev::k_scatter_sol<T><<<grid_rows64(n_), 256, 0, stream_>>>(X, d_x_, d_invp0_, n_, nrhs, d_scale_, 1);
  residual(d_b_, d_x_, d_r_, nrhs);
  colmax2(d_r_, d_x_, nrhs, d_red_);
  HIP_OK(hipGetLastError());
}

// z = S (L L^T)^{-1} S r with the FP32 factor of S A S (FP64 in and out).
void Engine::precond(const double* r, double* z, int nrhs) {
  float* Y = (float*)d_Y_;
  float* X = (float*)d_X_;
  dev::k_gather_rhs<float><<<grid_rows64(n_), 256, 0, stream_>>>(r, Y, d_invp0_, n_, nrhs, d_scale_);
  solve_sequence<float>(d_arena32_, Y, X, nrhs);
  dev::k_scatter_sol<float><<<grid_rows64(n_), 256, 0, stream_>>>(X, z, d_invp0_, n_, nrhs, d_scale_, 0);
  HIP_OK(hipGetLastError());
}

// One preconditioned-CG iteration for every right-hand side (FP64 vectors).
// Scalars live on the device: rho = d_s_[0,nrhs), p.q = [nrhs,2nrhs),
// rho_new = [2nrhs,3nrhs).
void Engine::pcg_iteration(int nrhs) {
  const int64_t cnt = (int64_t)n_ * nrhs;
  const unsigned g = grid_for(cnt);
  double* rho = d_s_;
  double* pq = d_s_ + nrhs;
  double* rho_new = d_s_ + 2 * nrhs;
  csr_mv(d_p_, nullptr, d_q_, nrhs);
  coldot(d_p_, d_q_, nrhs, pq);
  k_pcg_xr<<<g, 256, 0, stream_>>>(d_x_, d_r_, d_p_, d_q_, rho, pq, n_, nrhs);
  precond(d_r_, d_z_, nrhs);
  coldot(d_r_, d_z_, nrhs, rho_new);
  k_pcg_p<<<g, 256, 0, stream_>>>(d_p_, d_z_, rho_new, rho, n_, nrhs);
  dev::k_copy<<<1, 256, 0, stream_>>>(rho_new, rho, (int64_t)nrhs);
  colmax2(d_r_, d_x_, nrhs, d_red_);
  HIP_OK(hipGetLastError());
}

// (Re)start: r given; z = M^{-1} r; p = z; rho = r.z.
void Engine::pcg_restart(int nrhs) {
  const int64_t cnt = (int64_t)n_ * nrhs;
  precond(d_r_, d_z_, nrhs);
  dev::k_copy<<<grid_for(cnt), 256, 0, stream_>>>(d_z_, d_p_, cnt);
  coldot(d_r_, d_z_, nrhs, d_s_);
  HIP_OK(hipGetLastError());
}

// GMRES-based refinement for the FP32 static LDL^T factor of S A S
// (indefinite A): restarted GMRES(m) in FP64 on A x = b with the right
// preconditioner M^{-1} = S (L D L^T)^{-1} S, one Krylov space per
// right-hand side, batched over the right-hand sides.  Each step: z_k =
// M^{-1} v_k (stored), w = A z_k, classical Gram-Schmidt applied twice
// against v_0..v_k, v_{k+1} = w / ||w||; the (k+2) x (k+1) Hessenberg columns
// come to the host, where Givens rotations give the residual norm of every
// right-hand side.  A cycle ends when every estimate is below half the
// backward-error target (in the 2-norm, which bounds the infinity norm) or
// after m steps; then x += Z y and the true residual decides.  Fixed-point
// refinement x += M^{-1}(b - A x) needs the spectral radius of
// I - M^{-1} A below one, which the FP32 static LDL^T of an indefinite
// matrix does not guarantee; GMRES converges whenever the preconditioned
// matrix's spectrum is clustered, also with a few eigenvalues away from one
// (GMRES-IR, Carson and Higham 2017).
void Engine::gmres_ir(int nrhs, const std::vector<double>& bnorm, SolveReport& rep) {
  const int m = std::max(2, std::min(kGmresRestart, opt_.ir_max_iter));
  const int64_t cnt = (int64_t)n_ * nrhs;
  if (gm_cap_ < cnt * (2 * m + 1)) {
    dfree(d_gmV_);
    d_gmV_ = dalloc<double>((size_t)cnt * (2 * m + 1));  // V_0..V_m then Z_0..Z_{m-1}
    gm_cap_ = cnt * (2 * m + 1);
  }
  // h: the two Gram-Schmidt coefficient sets and the norm of step k,
  // (2 (k+1) + 1) x nrhs <= (2 m + 1) x nrhs; then the partial sums of up to
  // m vectors
  const int64_t hsize = (int64_t)(2 * m + 2) * nrhs;
  const int64_t need = hsize + (int64_t)kRedBlocks * m * nrhs;
  if (gm_hcap_ < need) {
    dfree(d_gmh_);
    gm_hcap_ = need;
    d_gmh_ = dalloc<double>((size_t)gm_hcap_);
  }
  double* V = d_gmV_;
  double* Z = d_gmV_ + cnt * (m + 1);
  double* h = d_gmh_;
  double* part = d_gmh_ + hsize;
  double* w = d_q_;
  const unsigned g = grid_for(cnt);
  std::vector<double> red(2 * (size_t)nrhs);
  auto berr_now = [&]() {
    residual(d_b_, d_x_, d_r_, nrhs);
    colmax2(d_r_, d_x_, nrhs, d_red_);
    HIP_OK(hipMemcpyAsync(red.data(), d_red_, red.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    double be = 0.0;
    for (int j = 0; j < nrhs; ++j) {
      const double den = normA_inf_ * red[2 * j + 1] + bnorm[j];
      be = std::max(be, den > 0 ? red[2 * j] / den : 0.0);
    }
    return be;
  };
  auto multidot = [&](int kv, const double* vec, double* out) {
    k_multidot_part<<<dim3(kRedBlocks, (unsigned)nrhs, (unsigned)kv), 256, 0, stream_>>>(V, cnt, vec, n_, nrhs, part);
    k_multidot_fin<<<(unsigned)(kv * nrhs), 256, 0, stream_>>>(part, out);
  };
  const int maxit = std::max(1, opt_.ir_max_iter);
  int total = 0;
  rep.converged = 0;
  // x = 0, r = b on entry
  std::vector<double> H((size_t)(m + 1) * m * nrhs), cs((size_t)m * nrhs), sn((size_t)m * nrhs),
      gv((size_t)(m + 1) * nrhs), hk((size_t)(m + 2) * nrhs * 2), y((size_t)m * nrhs), xnorm(nrhs, 0.0);
  const double sqrtn = std::sqrt((double)n_);
  bool first_cycle = true;
  // y = H^{-1} g per right-hand side (upper triangular kk x kk), uploaded to h
  auto solve_y = [&](int kk) {
    for (int j = 0; j < nrhs; ++j) {
      const double* Hc = &H[((size_t)j * (m + 1)) * m];
      for (int i = kk - 1; i >= 0; --i) {
        double t = gv[(size_t)i * nrhs + j];
        for (int c = i + 1; c < kk; ++c) t -= Hc[(size_t)c * (m + 1) + i] * y[(size_t)c * nrhs + j];
        const double d = Hc[(size_t)i * (m + 1) + i];
        y[(size_t)i * nrhs + j] = d != 0.0 ? t / d : 0.0;
      }
    }
    HIP_OK(hipMemcpyAsync(h, y.data(), (size_t)kk * nrhs * sizeof(double), hipMemcpyHostToDevice, stream_));
  };
  // true backward error of the candidate p = x + Z y after kk steps; on
  // success p becomes x.  Leaves r = b - A p (the cycle's basis does not
  // use r) and ||p||_inf in xnorm.
  double* pc = d_p_;
  auto check = [&](int kk) {
    solve_y(kk);
    dev::k_copy<<<g, 256, 0, stream_>>>(d_x_, pc, cnt);
    k_add_combo<<<g, 256, 0, stream_>>>(pc, Z, cnt, h, kk, n_, nrhs);
    residual(d_b_, pc, d_r_, nrhs);
    colmax2(d_r_, pc, nrhs, d_red_);
    HIP_OK(hipMemcpyAsync(red.data(), d_red_, red.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    double be = 0.0;
    for (int j = 0; j < nrhs; ++j) {
      xnorm[j] = std::max(xnorm[j], red[2 * j + 1]);
      const double den = normA_inf_ * red[2 * j + 1] + bnorm[j];
      be = std::max(be, den > 0 ? red[2 * j] / den : 0.0);
    }
    rep.backward_error = be;
    if (be > opt_.ir_tol) return false;
    dev::k_copy<<<g, 256, 0, stream_>>>(pc, d_x_, cnt);
    HIP_OK(hipGetLastError());
    rep.converged = 1;
    return true;
  };
  while (total < maxit) {
    // v_0 = r / ||r||
    coldot(d_r_, d_r_, nrhs, h);
    k_scale_cols<<<g, 256, 0, stream_>>>(V, d_r_, h, n_, nrhs);
    std::vector<double> beta2(nrhs);
    HIP_OK(hipMemcpyAsync(beta2.data(), h, nrhs * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    for (int j = 0; j < nrhs; ++j) {
      gv[j] = std::sqrt(std::max(0.0, beta2[j]));
      for (int i = 1; i <= m; ++i) gv[(size_t)i * nrhs + j] = 0.0;
      xnorm[j] = std::max(xnorm[j], red[2 * j + 1]);
    }
    int kk = 0;
    for (int k = 0; k < m && total < maxit; ++k) {
      double* vk = V + (int64_t)k * cnt;
      double* zk = Z + (int64_t)k * cnt;
      precond(vk, zk, nrhs);
      csr_mv(zk, nullptr, w, nrhs);
      // classical Gram-Schmidt, twice; the two coefficient sets are added
      multidot(k + 1, w, h);
      k_gs_update<<<g, 256, 0, stream_>>>(w, V, cnt, h, k + 1, n_, nrhs);
      multidot(k + 1, w, h + (int64_t)(k + 1) * nrhs);
      k_gs_update<<<g, 256, 0, stream_>>>(w, V, cnt, h + (int64_t)(k + 1) * nrhs, k + 1, n_, nrhs);
      double* nrm = h + (int64_t)2 * (k + 1) * nrhs;
      coldot(w, w, nrhs, nrm);
      k_scale_cols<<<g, 256, 0, stream_>>>(V + (int64_t)(k + 1) * cnt, w, nrm, n_, nrhs);
      HIP_OK(hipGetLastError());
      const size_t nh = (size_t)(2 * (k + 1) + 1) * nrhs;
      HIP_OK(hipMemcpyAsync(hk.data(), h, nh * sizeof(double), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      ++total;
      kk = k + 1;
      bool done = true, maybe = true;
      for (int j = 0; j < nrhs; ++j) {
        double* Hc = &H[((size_t)j * (m + 1)) * m];  // column-major (m+1) x m per right-hand side
        auto Hat = [&](int r, int c) -> double& { return Hc[(size_t)c * (m + 1) + r]; };
        for (int i = 0; i <= k; ++i) Hat(i, k) = hk[(size_t)i * nrhs + j] + hk[(size_t)(k + 1 + i) * nrhs + j];
        Hat(k + 1, k) = std::sqrt(std::max(0.0, hk[(size_t)2 * (k + 1) * nrhs + j]));
        for (int i = 0; i < k; ++i) {  // earlier rotations
          const double c = cs[(size_t)i * nrhs + j], s = sn[(size_t)i * nrhs + j];
          const double a = Hat(i, k), b = Hat(i + 1, k);
          Hat(i, k) = c * a + s * b;
          Hat(i + 1, k) = -s * a + c * b;
        }
        const double a = Hat(k, k), b = Hat(k + 1, k);
        const double rr = std::hypot(a, b);
        const double c = rr > 0 ? a / rr : 1.0, s = rr > 0 ? b / rr : 0.0;
        cs[(size_t)k * nrhs + j] = c;
        sn[(size_t)k * nrhs + j] = s;
        Hat(k, k) = rr;
        Hat(k + 1, k) = 0.0;
        const double gk = gv[(size_t)k * nrhs + j];
        gv[(size_t)k * nrhs + j] = c * gk;
        gv[(size_t)(k + 1) * nrhs + j] = -s * gk;
        const double est = std::fabs(gv[(size_t)(k + 1) * nrhs + j]);
        if (est > 0.5 * opt_.ir_tol * (normA_inf_ * xnorm[j] + bnorm[j])) done = false;
        if (est > sqrtn * opt_.ir_tol * (normA_inf_ * xnorm[j] + bnorm[j])) maybe = false;
      }
      if (done) break;
      // The estimates bound the 2-norm of the residual, the backward error
      // uses its infinity norm (||r||_inf >= ||r||_2 / sqrt(n)) and the
      // final ||x||_inf: where every estimate allows convergence with the
      // best known ||x||, the true backward error of x + Z y is computed (in
      // p; x stays) and a passing candidate ends the refinement.  The first
      // cycle checks after two steps, which also gives ||x||.
      if (maybe || (first_cycle && k == 1)) {
        if (check(kk)) {
          rep.iterations = total;
          return;
        }
      }
    }
    first_cycle = false;
    solve_y(kk);
    k_add_combo<<<g, 256, 0, stream_>>>(d_x_, Z, cnt, h, kk, n_, nrhs);
    HIP_OK(hipGetLastError());
    const double be = berr_now();  // also leaves r = b - A x for the next cycle
    for (int j = 0; j < nrhs; ++j) xnorm[j] = std::max(xnorm[j], red[2 * j + 1]);
    rep.iterations = total;
    rep.backward_error = be;
    if (be <= opt_.ir_tol) {
      rep.converged = 1;
      break;
    }
  }
}

SolveReport Engine::solve(const double* rhs, int nrhs, bool on_device) {
  if (!factored_) throw std::logic_error("Solve called without a successful Factorize");
  if (opt_.solve_dataflow)
    throw std::logic_error("the dependency-driven solve (solve_dataflow) is disabled in this build");
  if (nrhs <= 0) throw std::invalid_argument("nrhs must be positive");
  HIP_OK(hipSetDevice(opt_.device));
  auto t0 = std::chrono::steady_clock::now();
  ensure_solve_capacity(nrhs);
  if (opt_.precision == 0)
    ensure_ginv<double>(d_arena64_);
  else
    ensure_ginv<float>(d_arena32_);
  int* df_abort = use_df() ? d_df_state_ + 2 + 4 * (size_t)plan_.nsuper : nullptr;
  if (df_abort) dev::k_fill_zero<int><<<1, 64, 0, stream_>>>(df_abort, 1);
  const size_t cnt = (size_t)n_ * nrhs;
  HIP_OK(hipMemcpyAsync(d_b_, rhs, cnt * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  SolveReport rep;
  if (opt_.precision == 0) {
    double* Y = (double*)d_Y_;
    double* X = (double*)d_X_;
    run_graph(solve_graph(4 * (int64_t)nrhs), 4 * (int64_t)nrhs, [&] {
      dev::k_gather_rhs<double><<<grid_rows64(n_), 256, 0, stream_>>>(d_b_, Y, d_invp0_, n_, nrhs, nullptr);
      solve_sequence<double>(d_arena64_, Y, X, nrhs);
      dev::k_scatter_sol<double><<<grid_rows64(n_), 256, 0, stream_>>>(X, d_x_, d_invp0_, n_, nrhs, nullptr, 0);
      HIP_OK(hipGetLastError());
    });
  } else {
    // FP64 refinement around the FP32 factor of S A S (S = diag(s)).
    // SPD (LL^T): preconditioned conjugate gradients in FP64 with
    // M^{-1} = S (LL^T)^{-1} S, converging also where kappa(A) u_32 ~ 1;
    // static LDL^T (indefinite): restarted GMRES in FP64 with the same
    // preconditioner on the right (gmres_ir).
    // Both stop at the normwise backward error ir_tol, checked on the true
    // residual b - A x.
    std::vector<double> red(2 * (size_t)nrhs), bnorm(nrhs);
    auto read_berr = [&]() {
      HIP_OK(hipMemcpyAsync(red.data(), d_red_, red.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      double berr = 0.0;
      for (int j = 0; j < nrhs; ++j) {
        double den = normA_inf_ * red[2 * j + 1] + bnorm[j];
        berr = std::max(berr, den > 0 ? red[2 * j] / den : 0.0);
      }
      return berr;
    };
    dev::k_fill_zero<double><<<grid_for((int64_t)cnt), 256, 0, stream_>>>(d_x_, (int64_t)cnt);
    dev::k_copy<<<grid_for((int64_t)cnt), 256, 0, stream_>>>(d_b_, d_r_, (int64_t)cnt);
    colmax2(d_b_, d_x_, nrhs, d_red_);
    read_berr();
    for (int j = 0; j < nrhs; ++j) bnorm[j] = red[2 * j];
    rep.converged = 0;
    const int maxit = std::max(1, opt_.ir_max_iter);
    if (!opt_.ldl) {
      const int64_t kit = 4 * (int64_t)nrhs + 2, kst = 4 * (int64_t)nrhs + 3;
      run_graph(solve_graph(kst), kst, [&] { pcg_restart(nrhs); });
      for (int it = 0; it < maxit; ++it) {
        run_graph(solve_graph(kit), kit, [&] { pcg_iteration(nrhs); });
        double berr = read_berr();
        rep.iterations = it + 1;
        if (berr <= opt_.ir_tol || it + 1 == maxit) {
          residual(d_b_, d_x_, d_r_, nrhs);  // the true residual decides
          colmax2(d_r_, d_x_, nrhs, d_red_);
          berr = read_berr();
          rep.backward_error = berr;
          if (berr <= opt_.ir_tol) {
            rep.converged = 1;
            break;
          }
          run_graph(solve_graph(kst), kst, [&] { pcg_restart(nrhs); });
        }
        rep.backward_error = berr;
      }
    } else {
      gmres_ir(nrhs, bnorm, rep);
    }
  }
  int aborted = 0;
  if (df_abort) HIP_OK(hipMemcpyAsync(&aborted, df_abort, sizeof(int), hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
  collect_events();
  if (aborted)
    throw std::runtime_error("dependency-driven solve: a wait exceeded its budget (plan inconsistent)");
  last_solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  last_report = rep;
  return rep;
}

// ---------------------------------------------------------------------------
// Allocation-free FP64 solve.  The work vectors Y, X and Xf (n x nrhs each,
// row-major, permuted numbering) live in one file-scope device array of
// kNoAllocScratchBytes instead of buffers from hipMalloc, for callers that
// cannot allocate (code under graph capture, tight serving loops).  Every
// GPU holds its own copy of the array for the life of the process; its
// device address is looked up once per device (hipGetSymbolAddress with
// that device current) and only that address is passed to kernels.  Calls
// on one GPU share the array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::logic_error("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIP_OK(hipGetDevice(&cur));
    if (cur != device) HIP_OK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIP_OK(hipSetDevice(cur));
    HIP_OK(e);
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the solve scratch");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

void Engine::solve_noalloc(const double* d_rhs, double* d_x, int nrhs) {
  if (!factored_) throw std::logic_error("Solve called without a successful Factorize");
  if (nrhs <= 0) throw std::invalid_argument("nrhs must be positive");
  if (opt_.precision != 0)
    throw std::logic_error("the allocation-free solve runs FP64 factors only (precision FP64)");
  if (opt_.profile) throw std::logic_error("the allocation-free solve does not run in profile mode");
  const size_t cnt = (size_t)n_ * (size_t)nrhs;
  const size_t need = 3 * cnt * sizeof(double);
  if (need > kNoAllocScratchBytes)
    throw std::logic_error("allocation-free solve: 3 n nrhs doubles (" + std::to_string(need) +
                           " bytes) exceed the static scratch of " + std::to_string(kNoAllocScratchBytes) +
                           " bytes");
  HIP_OK(hipSetDevice(opt_.device));
  double* Y = (double*)noalloc_scratch(opt_.device);
  double* X = Y + cnt;
  double* Xf = X + cnt;
  auto t0 = std::chrono::steady_clock::now();
  ensure_ginv<double>(d_arena64_);
  dev::k_gather_rhs<double><<<grid_rows64(n_), 256, 0, stream_>>>(d_rhs, Y, d_invp0_, n_, nrhs, nullptr);
  solve_sequence<double>(d_arena64_, Y, X, nrhs, Xf);
  dev::k_scatter_sol<double><<<grid_rows64(n_), 256, 0, stream_>>>(X, d_x, d_invp0_, n_, nrhs, nullptr, 0);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(stream_));
  last_solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  last_report = SolveReport();
}

void Engine::get_solution(double* B, int nrhs, bool on_device) {
  if (nrhs > cap_nrhs_) throw std::logic_error("GetSolution: nrhs larger than the last Solve");
  HIP_OK(hipMemcpyAsync(B, d_x_, (size_t)n_ * nrhs * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
}

void Engine::get_panel(Int I, double* out) const {
  if (I < 0 || I >= plan_.nsuper) throw std::out_of_range("supernode index out of range");
  const SnInfo& s = plan_.sn[I];
  size_t cnt = (size_t)s.r * s.w;
  if (opt_.precision == 0) {
    HIP_OK(hipMemcpyAsync(out, d_arena64_ + s.off, cnt * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
  } else {
    std::vector<float> tmp(cnt);
    HIP_OK(hipMemcpyAsync(tmp.data(), d_arena32_ + s.off, cnt * sizeof(float), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < cnt; ++i) out[i] = tmp[i];
  }
}

}  // namespace spk
