/*
 * C ABI of the MI355X-native supernodal Cholesky / static LDL^T path.
 *
 * Every entry point returns an int status: 0 on success, < 0 for an argument
 * or device error (spk_last_error() holds the message), and for the
 * factorization > 0 = first failing pivot column (1-based, in the permuted
 * order; LAPACK "info" semantics) where symPACK throws
 * "Matrix is not HPD." (src/lapack.cpp:293-306).
 *
 * The life cycle mirrors symPACKMatrix2D (include/sympack/symPACKMatrix2D.hpp
 * :2612-2627, driver/run_sympack2D.cpp:74-128):
 *   spk_create -> spk_init -> spk_symbolic_factorization ->
 *   spk_distribute_matrix -> spk_factorize -> spk_solve -> spk_get_solution
 *   -> spk_destroy
 * Matrices are given as the lower triangle (diagonal included) in compressed
 * sparse column form, as DistSparseMatrix<T> holds them on one rank
 * (include/sympack/DistSparseMatrix.hpp:46-61).  Right-hand sides and
 * solutions are column-major n x nrhs in the caller's numbering.
 */
#ifndef SYMPACK_AMD_H_
#define SYMPACK_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ordering codes: numeric values of symPACK::OrderingType
 * (include/sympack/Types.hpp:101). */
#define SPK_ORDER_NATURAL 0
#define SPK_ORDER_RCM 1
#define SPK_ORDER_MMD 2
#define SPK_ORDER_AMD 3
#define SPK_ORDER_NDBOX 4
#define SPK_ORDER_NDGRID 5
#define SPK_ORDER_USER 10

/* symPACK::DecompositionType (include/sympack/Types.hpp:97). */
#define SPK_FACT_LL 0
#define SPK_FACT_LDL 1

/* Arithmetic of the factor. */
#define SPK_PREC_FP64 0        /* FP64 factor and solve (the drop-in) */
#define SPK_PREC_FP32_IR 1     /* FP32 factor, FP64 iterative refinement */

typedef struct spk_matrix_s* spk_handle;

/* Options: the fields of symPACKOptions that run_sympack2D sets
 * (include/sympack/Types.hpp:105-188, driver/utils.hpp:227-452). */
typedef struct spk_options {
  int ordering;       /* SPK_ORDER_*; default MMD */
  int decomposition;  /* SPK_FACT_*; default LL */
  int nrelax0, nrelax1, nrelax2; /* RelaxationParameters; default 8, 32, 64 */
  int maxsize;        /* maximum supernode width, 0 = unlimited (driver default) */
  double zrelax0, zrelax1, zrelax2; /* 0.8, 0.1, 0.05 */
  int precision;      /* SPK_PREC_*; default FP64 */
  int ir_max_iter;    /* refinement steps for SPK_PREC_FP32_IR (default 30): preconditioned
                         conjugate gradient iterations for LL^T, GMRES steps (restarted every
                         20) for static LDL^T */
  double ir_tol;      /* normwise backward-error target (default 1e-12) */
  const int* perm;    /* SPK_ORDER_USER: perm[new] = old, any base (symPACKOptions.perm) */
  int verbose;
  int profile;        /* time every kernel launch with HIP events */
  int device;         /* HIP device ordinal */
  int merge_updates;  /* 1 (default): a chain child's update outside its parent's columns is
                         computed inside the parent's Schur update (one scatter instead of two)
                         on etree levels with enough update tiles to fill the GPU; 2: on every
                         level; 0: every supernode scatters its whole update */
  int solve_dataflow; /* must be 0 (default): one launch per etree level and kernel kind.
                         The dependency-driven sweeps (one persistent launch each) are
                         disabled in this build; Solve rejects 1 */
  int refine;         /* supernode reordering (symPACKOptions.order_refinement_str,
                         symPACKMatrix2D.hpp:3090-3263): 0 = NO (default), or
                         10*ordflag + altflag: 10, 11, 20, 21, 30, 31 for SET10 .. SET31
                         (SET = 31), 40 = TSPB (traveling-salesman tour with skeleton
                         compression), 41 = any other TSPB* name (tour without it) */
  int solve_inverse;  /* 1 (default): where every supernode is at most 160 columns wide, the
                         solve uses the partitioned inverse G_I = [I; L_off] L_II^{-1} of each
                         supernode (prepared once per factorization), one launch per etree level
                         and sweep; for LL^T and for the FP32 factor under refinement, not for
                         FP64 static LDL^T (explicit inverses of its indefinite diagonal blocks
                         lose accuracy).  0: block substitution, four launches per level */
  int autotune;       /* 1: the first spk_factorize of a structure with at least 1e9
                         Schur-update flops times every etree level's Schur launch under each
                         kernel (64 x 64 tiles staged through LDS, 128 x 64 tiles, 64 x 64 tiles
                         with one wave per tile), task order (plan / longest first) and k split
                         (1-8 workgroups per tile) on the device and keeps the fastest
                         (spk_get_autotune), for structures factored many times (tuning costs
                         tens of factorizations once); 0 (default): the plan's heuristic.  No
                         reference counterpart */
  int schur_kernel;   /* 0 (default): the Schur-update kernel the autotuner (or, untuned, the
                         plan) chooses per level; 1: 64 x 64 tiles staged through LDS, 2: 128 x
                         64 tiles, 3: 64 x 64 tiles with one wave per tile, on every level (no
                         tuning); other values are rejected by spk_init */
} spk_options;

typedef struct spk_stats {
  int32_t n, nsuper, nlevels, max_width;
  double flops;           /* sum cc_j^2 over unrelaxed column counts (the "Flops:" line) */
  int64_t nnz_l;          /* sum cc_j (the "NNZ in L factor:" line) */
  int64_t factor_entries; /* stored factor entries, relaxed supernodes incl. upper diag blocks */
  double schur_flops;     /* sum 2 m n k of the Schur updates */
  double inpanel_flops, diag_flops;
  int64_t schur_entries;  /* scattered update entries */
  int64_t launches_factor, launches_solve;
  double factor_ms, solve_ms;     /* wall time of the last Factorize / Solve */
  int ir_iterations;              /* last Solve */
  double backward_error;          /* last Solve (FP32+IR), -1 if not computed */
  double fp32_shift;              /* diagonal shift of the equilibrated matrix in the last
                                     FP32 factor (0 unless its Cholesky broke down) */
  int ir_converged;               /* last Solve: 1 if the refinement reached ir_tol (always 1
                                     for FP64), 0 if it stopped after ir_max_iter steps */
} spk_stats;

/* Per-kernel-kind event timings (profile mode).  kind: 0 diagonal factor,
 * 7 panel TRSM, 9 fused diagonal factor + panel TRSM, 1 in-panel update,
 * 2 Schur update, 3 forward diagonal, 4 forward off-diagonal, 5 backward
 * off-diagonal, 6 backward diagonal, 10 forward / 11 backward
 * partitioned-inverse solve, 12 its preparation. */
int spk_get_kernel_stats(spk_handle h, int kind, int64_t* launches, double* ms);
int spk_reset_kernel_stats(spk_handle h);
/* Switch per-launch event timing on/off after spk_init (no reference
 * counterpart; the reference's timing is the SPROFILE build, timer.hpp:56-85). */
int spk_set_profile(spk_handle h, int on);
/* As spk_set_profile, timing only the kernel kinds whose bit is set in mask. */
int spk_set_profile_mask(spk_handle h, int on, unsigned mask);
/* HBM footprint of the handle after spk_distribute_matrix (plan, maps,
 * values, factor) plus solve buffers for nrhs right-hand sides; valid after
 * spk_symbolic_factorization.  Used for admission by memory budget. */
int spk_device_bytes(spk_handle h, int nrhs, int64_t* bytes);
/* Free the handle's device memory, keeping the symbolic analysis; the values
 * must be distributed again before the next spk_factorize. */
int spk_release_device(spk_handle h);
/* Replay factor and solve launch sequences from captured HIP graphs (on = 1,
 * the default) or launch each kernel from the host (on = 0).  No reference
 * counterpart (the reference has no device launch sequences). */
int spk_set_graphs(spk_handle h, int on);
/* Per-launch records of the profiled calls since the last reset: kind, grid
 * size (workgroups) and event time; returns the number of records (<= max). */
int spk_get_launch_trace(spk_handle h, int32_t* kinds, int64_t* grids, double* ms, int max);
/* Etree level (height above the leaves) of every record of
 * spk_get_launch_trace, -1 for launches outside the level sequences; returns
 * the number of records (<= max).  No reference counterpart. */
int spk_get_launch_levels(spk_handle h, int32_t* levels, int max);
/* The autotuner's choice per Schur launch (one per etree level with updates)
 * after the first factorization: level, k split (workgroups per tile), launch
 * variant (bit 0: tasks longest first instead of plan order; bit 1: 128 x 64
 * output tiles; bit 2: 64 x 64 tiles with one wave per tile; neither: 64 x 64
 * tiles staged through LDS), best time of the default configuration
 * and of the chosen one (ms); returns the number of records (<= max), 0 when
 * nothing was tuned.  No reference counterpart. */
int spk_get_autotune(spk_handle h, int32_t* level, int32_t* split, int32_t* order, double* ms_default,
                     double* ms_best, int max);
/* Etree level of every supernode (0-based supernode index, 0 = leaves): the
 * level whose launches factor it and solve with it.  No reference
 * counterpart (the reference schedules tasks, not levels). */
int spk_get_supernode_levels(spk_handle h, int32_t* level /* nsuper */);

/* Library set-up, as the reference's C entry points (include/sympack.hpp:31-36,
 * src/global.cpp:114-185).  One rank: symPACK_Rank always returns 0. */
int symPACK_Init(int* argc, char*** argv);
int symPACK_Finalize(void);
int symPACK_Rank(int* rank);
/* Replaces symPACK_cuda_setup (src/global.cpp:32-60): select the HIP device. */
int symPACK_device_setup(int device);

void spk_default_options(spk_options* opt);
/* Parse run_sympack2D flags (-ordering, -fact, -relax a b c, -b, -v,
 * -precision, -ir_tol, -ir_iter) into opt (driver/utils.hpp:227-452).
 * Unknown flags are ignored, as the reference's option map does. */
int spk_options_from_args(spk_options* opt, int argc, const char* const* argv);

int spk_create(spk_handle* h);
int spk_destroy(spk_handle h);
/* symPACKMatrix2D::Init (symPACKMatrix2D.hpp:2649-2695) */
int spk_init(spk_handle h, const spk_options* opt);
/* symPACKMatrix2D::SymbolicFactorization (symPACKMatrix2D.hpp:2699-6334):
 * lower-triangular CSC pattern with base `baseval` (0 or 1). */
int spk_symbolic_factorization(spk_handle h, int32_t n, const int64_t* colptr,
                               const int32_t* rowind, int baseval);
/* symPACKMatrix2D::DistributeMatrix (symPACKMatrix2D.hpp:6337-6576):
 * values of the same CSC, host memory / device memory. */
int spk_distribute_matrix(spk_handle h, const double* nzval);
int spk_distribute_matrix_device(spk_handle h, const double* d_nzval);
/* symPACKMatrix2D::Factorize (symPACKMatrix2D.hpp:6579-6639). */
int spk_factorize(spk_handle h);
/* symPACKMatrix2D::Solve (symPACKMatrix2D.hpp:6643-6710): RHS is borrowed and
 * not modified; rhs_size is unused, as in the reference's 2D path. */
int spk_solve(spk_handle h, const double* rhs, int nrhs, int64_t rhs_size);
int spk_solve_device(spk_handle h, const double* d_rhs, int nrhs);
/* Allocation-free solve: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops): Solve + GetSolution of a factored FP64 handle from device-resident
 * d_rhs into device-resident d_x (both column-major n x nrhs), with the work
 * vectors in a static per-device array of 128 MiB instead of hipMalloc'd
 * buffers, kernels launched directly (no graph instantiation).  Limit:
 * 3 n nrhs doubles <= 128 MiB (n nrhs <= 5592405); larger requests, an
 * FP32_IR handle or profile mode return -2.  Calls that overlap on one GPU
 * share the array and race. */
int spk_solve_device_noalloc(spk_handle h, const double* d_rhs, double* d_x, int nrhs);
/* symPACKMatrix2D::GetSolution (symPACKMatrix2D.hpp:6713-6757). */
int spk_get_solution(spk_handle h, double* B, int nrhs);
int spk_get_solution_device(spk_handle h, double* d_B, int nrhs);

/* Accessors (symPACKMatrixBase.hpp:56-103): all arrays 1-based as in symPACK. */
int spk_get_stats(spk_handle h, spk_stats* st);
int spk_get_supernodal_partition(spk_handle h, int32_t* xsuper /* nsuper+1 */);
int spk_get_sup_membership(spk_handle h, int32_t* supmem /* n */);
int spk_get_ordering(spk_handle h, int32_t* perm /* n */, int32_t* invp /* n */);
int spk_get_etree(spk_handle h, int32_t* parent /* n, postordered */);
int spk_get_column_counts(spk_handle h, int32_t* cc_unrelaxed, int32_t* cc_relaxed);
int spk_get_structure_size(spk_handle h, int64_t* nnz_lindx);
int spk_get_structure(spk_handle h, int64_t* xlindx /* nsuper+1 */, int32_t* lindx);
/* 2D cell map (symPACKMatrix2D.hpp:3317-3396): sizes, then arrays. */
int spk_get_cell_counts(spk_handle h, int64_t* ncells, int64_t* nblocks);
int spk_get_cells(spk_handle h, int32_t* cell_K, int32_t* cell_I, int32_t* cell_nrows,
                  int32_t* cell_nblocks, int32_t* block_first_row, int32_t* block_nrows);
/* Schur-update merging chosen by the plan (no counterpart in the reference,
 * whose tasks update each target cell from one source supernode):
 * deferred_into[I] = 0-based parent that computes the part of supernode I's
 * update outside the parent's columns, -1 where I scatters its whole update. */
int spk_get_update_merges(spk_handle h, int32_t* deferred_into /* nsuper */);
/* Checks the task lists of the dependency-driven solve (no counterpart in
 * the reference): 0 when a single worker taking tasks in ticket order never
 * waits and every counter drains, < 0 with spk_last_error() otherwise. */
int spk_validate_solve_plan(spk_handle h);
/* Factor panel of supernode I (1-based): r x w column-major, structure rows
 * by supernode columns; LDL panels hold D on the diagonal. */
int spk_get_panel_dims(spk_handle h, int32_t I, int32_t* nrows, int32_t* ncols);
int spk_get_factor_panel(spk_handle h, int32_t I, double* out);

/* The ordering routine alone, as symPACK::Ordering::{MMD,RCM,AMD,NDBOX,NDGRID}
 * computes it (src/Ordering.cpp:103-525) before the elimination-tree
 * postorder: graph = lower CSC pattern expanded, diagonal dropped.
 * perm[new-1] = old, invp[old-1] = new (1-based). */
int spk_compute_ordering(int ordering, int32_t n, const int64_t* colptr, const int32_t* rowind,
                         int baseval, int32_t* perm, int32_t* invp);

/* Rutherford-Boeing / Harwell-Boeing I/O (ReadHB_PARA, utility.hpp:1274-1482). */
int spk_read_rb(const char* filename, int32_t* n, int64_t* nnz);
int spk_read_rb_data(const char* filename, int64_t* colptr, int32_t* rowind, double* nzval);
int spk_write_rb(const char* filename, int32_t n, const int64_t* colptr, const int32_t* rowind,
                 const double* nzval, const char* title);

/* ReadMatrix (utility.hpp:1693-1716) on one rank: format "CSC" (binary,
 * ParaReadDistSparseMatrix, utility.hpp:1037-1265), "HB" / "RB" /
 * "HARWELL_BOEING" (ReadHB_PARA) or "matrix" (text,
 * ReadDistSparseMatrixFormatted, utility.hpp:1486-1688); any other name is
 * "Unknown matrix format." (status -2).  Returns the lower triangle, 1-based,
 * rows sorted per column for CSC and HB/RB and in file order for "matrix"
 * (a full symmetric "matrix" file is reduced to its lower triangle). */
int spk_read_matrix(const char* filename, const char* format, int32_t* n, int64_t* nnz);
int spk_read_matrix_data(const char* filename, const char* format, int64_t* colptr, int32_t* rowind,
                         double* nzval);
/* Writes a lower-triangular CSC (any base) in one of the formats above (no
 * reference counterpart: the reference only reads). */
int spk_write_matrix(const char* filename, const char* format, int32_t n, const int64_t* colptr,
                     const int32_t* rowind, const double* nzval);

const char* spk_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SYMPACK_AMD_H_ */
