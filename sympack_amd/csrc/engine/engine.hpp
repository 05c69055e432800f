// Device engine: owns the HBM-resident factor, the plan and the launch
// sequences for factorization and solve on one GPU.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../analysis/analysis.hpp"
#include "plan.hpp"

namespace spk {

constexpr int kRedBlocks = 64;  // blocks per column of the refinement reductions
constexpr int kGmresRestart = 20;  // GMRES restart length of the LDL^T refinement
// partitioned-inverse sweeps: workgroups a level launch aims at before it
// splits its right-hand-side blocks over the grid
constexpr int64_t kGinvGrid = 1024;
// Size of the per-device array behind Engine::solve_noalloc: Y, X and Xf,
// 3 n nrhs doubles, at most 128 MiB (n nrhs <= 5592405).
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
// Device address of that array on `device` (looked up once per device).
void* noalloc_scratch(int device);

struct EngineOptions {
  int ordering = ORD_MMD;
  int ldl = 0;            // 0: LL^T, 1: static LDL^T
  RelaxParams relax;
  int precision = 0;      // 0: FP64 factor/solve; 1: FP32 factor + FP64 refinement
  int ir_max_iter = 30;
  double ir_tol = 1e-12;  // normwise backward-error target of the refinement
  int device = 0;
  int profile = 0;        // record HIP events around every launch
  int verbose = 0;
  int merge_updates = 1;  // 0 off, 1 on levels that fill the GPU, 2 every qualifying chain child
  int solve_dataflow = 0; // dependency-driven solve sweeps: disabled, requesting it is an error
  int solve_inverse = 1;  // partitioned-inverse solve where every supernode is <= kStep wide
  int refine = 0;         // 10*ordflag + altflag of the supernode reordering, 0 = none
  int autotune = 1;       // time the Schur launch configurations of every level on the first
                          // factorization of a plan with at least kAutotuneMinFlops Schur flops
  int schur_kernel = 0;   // 0: tuner / plan default; 1, 2, 3: k_update, k_schur_wide, k_schur_w1
                          // on every level (no tuning when set)
  std::vector<Int> user_perm;
};

// profile mask bit per kernel kind (LaunchRange::kind 0..12)
constexpr unsigned kAllKinds = (1u << 13) - 1;

struct KindStats {
  int64_t launches = 0;
  double ms = 0.0;        // summed event time (profile mode)
};

// One Schur launch range as the autotuner left it: the default
// configuration (plan order, heuristic k split) and the fastest one measured.
struct TuneRecord {
  int level;
  int64_t tiles;
  int split0, alt0;
  double ms0;         // best time of the default configuration
  int split, alt;     // chosen configuration
  double ms;          // its best time
  int candidates;     // configurations timed
};
// plans with fewer Schur flops are not tuned (their factorization takes
// about as long as tuning it)
constexpr double kAutotuneMinFlops = 1e9;

struct SolveReport {
  int iterations = 0;       // refinement steps taken (0 for plain FP64)
  double backward_error = -1.0;  // max over RHS of ||b-Ax||inf/(||A||inf||x||inf+||b||inf)
  int converged = 1;
};

class Engine {
 public:
  Engine();
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  void init(const EngineOptions& o);
  void set_user_perm(const std::vector<Int>& perm) { opt_.user_perm = perm; }
  void set_profile(bool on, unsigned mask = kAllKinds) {
    if ((on ? 1 : 0) != opt_.profile || mask != profile_mask_) destroy_graphs();
    opt_.profile = on ? 1 : 0;
    profile_mask_ = mask;
  }
  // Replay the factor and solve launch sequences from captured HIP graphs
  // (default) or launch every kernel from the host.
  void set_graphs(bool on) {
    if (!on) destroy_graphs();
    use_graphs_ = on;
  }
  void symbolic(Int n, const Ptr* colptr, const Idx* rowind, int baseval);
  void distribute(const double* nzval, bool on_device);
  int factorize();
  SolveReport solve(const double* rhs, int nrhs, bool on_device);
  void get_solution(double* B, int nrhs, bool on_device);
  // FP64 solve of device-resident b into device-resident x that allocates
  // nothing: work vectors in the static per-device scratch, kernels launched
  // directly (no graph instantiation).  Rejects precision FP32_IR, profile
  // mode and 3 n nrhs doubles beyond kNoAllocScratchBytes.
  void solve_noalloc(const double* d_rhs, double* d_x, int nrhs);

  // HBM bytes this engine holds after DistributeMatrix (plan, maps, values,
  // factor) plus the solve buffers for nrhs right-hand sides.
  int64_t device_bytes(int nrhs) const;
  // Frees every device buffer; the symbolic analysis is kept, the values
  // must be distributed again before the next factorization.
  void release_device();
  const Symbolic& sym() const { return sym_; }
  const Plan& plan() const { return plan_; }
  const EngineOptions& options() const { return opt_; }
  hipStream_t stream() const { return stream_; }
  // copy the panel of supernode I (0-based) to host, as double
  void get_panel(Int I, double* out) const;
  std::vector<KindStats> kind_stats;  // indexed by LaunchRange::kind
  struct LaunchRecord {
    int kind;
    int64_t grid;
    double ms;
    int level;  // etree level of the launch (-1: outside the level sequences)
  };
  std::vector<LaunchRecord> trace;    // per launch of the last profiled calls
  std::vector<TuneRecord> tune_log;   // per Schur launch range, after autotuning
  bool tuned() const { return tuned_; }
  void reset_stats();
  double last_factor_ms = 0.0, last_solve_ms = 0.0;
  SolveReport last_report;
  bool factored() const { return factored_; }

 private:
  void free_device();
  template <typename T>
  void factor_sequence(T* arena);
  template <typename T>
  void launch_schur(const LaunchRange& r, int alt, int ks, T* arena, hipStream_t st);
  // Times every Schur launch range under each candidate configuration (task
  // order x k split) on the arena before the factorization assembles it, and
  // keeps the fastest in plan_.factor_seq.
  template <typename T>
  void autotune_schur(T* arena);
  bool tuned_ = false;
  template <typename T>
  void launch_diag(int bucket, bool ldl, unsigned g, const Task4* t, T* arena);
  template <typename T, bool BWD>
  void launch_sdiag(int bucket, unsigned g, const Task4* t, const T* arena, const T* Y, T* X, int nrhs);
  template <typename T>
  void launch_ptrsm(int bucket, bool ldl, unsigned g, const Task4* t, T* arena);
  template <typename T>
  void launch_diagtrsm(int bucket, bool ldl, unsigned g, const Task4* t, T* arena);
  template <typename T>
  void solve_sequence(const T* arena, T* Y, T* X, int nrhs, T* Xf = nullptr);
  template <typename T>
  void ginv_sequence(T* Y, T* X, T* Xf, int nrhs);
  // the partitioned-inverse solve applies to this plan and is switched on.
  // Not for FP64 static LDL^T: explicit inverses of its unit-lower diagonal
  // blocks (indefinite, static pivots) raise the backward error several
  // times over block substitution (80^3-like shifted Laplacians: 6x on the
  // host in exact block inverses, 26x measured at 48^3), beyond the
  // drop-in's 10x bar; as the FP32 preconditioner under refinement it is
  // used, the refinement restores the accuracy
  bool use_ginv() const {
    return plan_.ginv_ok && opt_.solve_inverse != 0 && (opt_.ldl == 0 || opt_.precision == 1);
  }
  // prepares G from the current factor once per factorization
  template <typename T>
  void ensure_ginv(const T* arena);
  template <typename T>
  void solve_sequence_df(const T* arena, T* Y, T* X, int nrhs);
  // The dependency-driven sweeps did not finish on MI355X (waits on counters
  // updated from other XCDs), so they are switched off: see solve().
  bool use_df() const { return false; }
  template <typename T>
  void assemble(T* arena);
  void residual(const double* b, const double* x, double* r, int nrhs);
  void csr_mv(const double* x, const double* b, double* out, int nrhs);
  void colmax2(const double* r, const double* x, int nrhs, double* out);
  void coldot(const double* a, const double* b, int nrhs, double* out);
  void precond(const double* r, double* z, int nrhs);
  void pcg_iteration(int nrhs);
  void pcg_restart(int nrhs);
  void gmres_ir(int nrhs, const std::vector<double>& bnorm, SolveReport& rep);
  void launch_begin(int kind, int64_t grid = 0, hipStream_t st = nullptr, int level = -1);
  void launch_end(int kind, hipStream_t st = nullptr);

  // captured launch sequences
  struct Graph {
    hipGraphExec_t exec = nullptr;
    int64_t key = -1;
    uint64_t used = 0;
    std::vector<int> kinds;  // kernel kind of every captured launch (for the counters)
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;  // profile mode
    std::vector<int64_t> grids;
  };
  template <typename F>
  void run_graph(Graph& g, int64_t key, F&& body);
  Graph& solve_graph(int64_t key);
  void destroy_graph(Graph& g);
  void destroy_graphs();
  bool use_graphs_ = true;
  bool capturing_ = false;
  Graph fgraph_;
  std::vector<Graph> sgraphs_;
  uint64_t graph_clock_ = 0;
  std::vector<Graph*> pending_graphs_;  // launched since the last collect_events
  std::vector<int> cap_kinds_;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> cap_events_;
  std::vector<int64_t> cap_grids_;
  template <typename T>
  void ir_step(T* Y, T* X, int nrhs);

  EngineOptions opt_;
  Symbolic sym_;
  Plan plan_;
  bool have_symbolic_ = false, have_values_ = false, factored_ = false, device_ready_ = false;
  void ensure_device();
  hipStream_t stream_ = nullptr;

  // caller CSC (0-based) kept for assembly and residuals
  Int n_ = 0;
  int64_t nnzA_ = 0;
  std::vector<int64_t> h_colptr_;
  std::vector<int32_t> h_rowind_;
  // device memory
  void* d_sn_ = nullptr;
  int32_t* d_lindx0_ = nullptr;
  int64_t* d_tcol_ = nullptr;
  int64_t* d_tmap_ = nullptr;
  int32_t* d_rel_ = nullptr;
  MergeInfo* d_merges_ = nullptr;
  int32_t* d_mmap_ = nullptr;
  int64_t* d_amap_ = nullptr;
  int32_t* d_invp0_ = nullptr;  // original -> permuted row, 0-based
  void* d_tasks_[12] = {nullptr};  // diag, trsm, inpanel, schur, sdiag, soff, soffm, fused, partitioned-inverse tiles,
                                   // schur longest first, schur 128 x 64 (plan order, longest first)
  int* d_info_ = nullptr;
  int* d_fcnt_ = nullptr;  // per supernode: loads of the fused block factor + TRSM
  // dependency-driven solve
  Task4* d_df_fwd_ = nullptr;
  Task4* d_df_bwd_ = nullptr;
  int32_t* d_df_fend_ = nullptr;
  int32_t* d_df_bend_ = nullptr;
  int32_t* d_df_tgt_ = nullptr;
  int32_t* d_df_dcnt0_ = nullptr;
  int32_t* d_df_rcnt0_ = nullptr;
  int64_t* d_invb_ = nullptr;
  int* d_df_state_ = nullptr;  // head[2], dcnt, dflag, rcnt, bflag, abort
  int df_grid_ = 0;
  int df_nf_ = 0, df_nb_ = 0;  // task counts of the uploaded dataflow lists
  int64_t* d_acolptr_ = nullptr;
  int32_t* d_arowind_ = nullptr;
  int32_t* d_acolidx_ = nullptr;
  double* d_vals_ = nullptr;
  double* d_scale_ = nullptr;  // equilibration s_i (FP32 mode)
  double* d_arena64_ = nullptr;
  void* d_dinv_ = nullptr;     // inverted 16x16 diagonal tiles of the factor
  void* d_garena_ = nullptr;   // partitioned-inverse solve factor G (panel shapes, factor precision)
  bool ginv_ready_ = false;    // G holds the current factor's inverse
  float* d_arena32_ = nullptr;
  // solve work
  int cap_nrhs_ = 0;
  void* d_Y_ = nullptr;
  void* d_X_ = nullptr;
  void* d_Xf_ = nullptr;  // forward-sweep results of the partitioned-inverse solve
  double* d_b_ = nullptr;
  double* d_x_ = nullptr;
  double* d_r_ = nullptr;
  double* d_red_ = nullptr;
  // A as full CSR (both triangles, original numbering) for the refinement
  // residuals and CG products (FP32 factor mode): row pointers, columns,
  // source entry in the caller's lower CSC, values gathered from it
  int64_t* d_crow_ = nullptr;
  int32_t* d_ccol_ = nullptr;
  int64_t* d_csrc_ = nullptr;
  double* d_cval_ = nullptr;
  int64_t nnz_full_ = 0;
  double* d_part_ = nullptr;  // per-block partial reductions, nrhs x kRedBlocks x 2
  double* d_z_ = nullptr;  // preconditioned CG vectors (FP32 factor mode)
  double* d_p_ = nullptr;
  double* d_q_ = nullptr;
  double* d_s_ = nullptr;  // per-RHS CG scalars
  double* d_gmV_ = nullptr;  // GMRES basis V_0..V_m and preconditioned vectors Z_0..Z_{m-1}
  double* d_gmh_ = nullptr;  // GMRES dot products, norms and partial sums
  int64_t gm_cap_ = 0, gm_hcap_ = 0;
 public:
  double fp32_shift_ = 0.0;  // diagonal shift of S A S used by the last FP32 factor
 private:
  int fp32_level_ = 0;
  void ensure_solve_capacity(int nrhs);
  double normA_inf_ = 0.0;
  // profiling events
  std::vector<hipEvent_t> ev_pool_;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open_;
  std::vector<int64_t> ev_grid_;
  std::vector<int> ev_level_;
  unsigned profile_mask_ = kAllKinds;  // kinds whose launches get events
  size_t ev_next_ = 0;
  void collect_events();
};

std::string hip_error_string(hipError_t e);

}  // namespace spk
