// Device execution plan for the supernodal factorization and solve.
//
// The plan is built once per symbolic structure on the host and uploaded to
// HBM.  It fixes:
//  * the factor layout: one contiguous arena holding every supernode's panel
//    (all rows of its structure x its columns), column-major with leading
//    dimension = number of structure rows, so a column of a panel is a
//    contiguous run of HBM;
//  * relative-index maps, so the Schur-update kernel scatters into ancestor
//    panels without searching block lists;
//  * the assembly map from the caller's CSC entries to arena offsets;
//  * etree-level task lists for every batched kernel launch.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../analysis/analysis.hpp"

namespace spk {

constexpr int kNB = 64;      // diagonal block width of the blocked solve
constexpr int kTile = 64;    // update tile edge (rows x cols)
constexpr int kStep = 160;   // panel columns factored per step (diagonal block <= 160 x 160)
constexpr int kSub = 16;     // sub-block edge inside a diagonal block (MFMA 16x16 tiles)
constexpr int kTrsmRows = 64;  // panel rows per fused block factor + TRSM workgroup (16 per wave)
constexpr int kTrsmRowsWide = 128;  // panel rows per k_ptrsm workgroup (8 waves of 16 rows)
constexpr int kSchurKC = 32;   // k columns per LDS chunk of the Schur update
// Schur launches with fewer tasks than this (one per CU of the 256) split
// each task's k chunks over several workgroups; more splitting costs more in
// extra scatter than it gains in latency
constexpr int64_t kSplitTarget = 256;

// LDS bucket of a diagonal block with nb 16-column sub-blocks: kernels are
// compiled for nb <= 1, 2, 4, 6, 10.
inline int diag_bucket(int nb) { return nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : nb <= 6 ? 6 : 10; }

struct SnInfo {            // mirrored on the device (16-byte aligned)
  int64_t off;             // panel offset in the arena (elements)
  int64_t upd;             // offset into tcol/tmap for the panel's off-diagonal rows
  int32_t r;               // structure rows (= panel leading dimension)
  int32_t w;               // width (columns)
  int32_t fc;              // first column, 0-based, permuted numbering
  int32_t lx;              // offset of the structure in lindx0
};

struct Task4 {
  int32_t a, b, c, d;
};

// A child supernode C whose Schur update is split in two: the part landing in
// its parent P's columns is applied at C's level; the rest (rows and columns
// both in struct(P)) is deferred and computed inside P's Schur update as
// extra k columns, so those target entries are scattered once instead of
// twice.  mmap[moff + a] = row of C's panel holding P's off-diagonal row a
// (-1 where C has no such row).
struct MergeInfo {         // mirrored on the device
  int64_t off;             // C's panel offset
  int64_t moff;            // offset of the row map in mmap
  int32_t ld;              // C's panel leading dimension (rows)
  int32_t w;               // C's width (extra k columns)
  int32_t child;           // C
  int32_t pad;
};

// One run of consecutive off-diagonal rows [gb, ge) of supernode I whose
// rows are columns of the same ancestor J.  The relative indices of all of
// I's off-diagonal rows a >= gb inside struct(J) are
// rel[base + a - gb] (a in [gb, m_I)); they, and the scatter bases tcol /
// tmap of the rows [gb, ge), are filled on the device from this table
// (binary search of each row in struct(J)).
struct RelGroup {          // mirrored on the device
  int64_t base;
  int32_t I, J, gb, ge;
};

// Host and device code (the plan is counted on the host, the task lists are
// written on the device by the same formulas).
#if defined(__HIPCC__)
#define SPK_HD __host__ __device__
#else
#define SPK_HD
#endif

// Schur output tiles of one supernode with m off-diagonal rows: 64 x 64
// tiles (ti, tj), tj <= ti, tj < ntj (ntj = the column-tile count, fewer for
// a deferred supernode's column cut), in order ti then tj; 128 x 64 tiles
// (ti, tj), ti < ceil(m / 128), tj <= 2 ti + 1, tj < ntj.  *_before(ti, ntj):
// tiles of the rows before row tile ti.
SPK_HD inline int64_t schur_tiles_before(int64_t ti, int64_t ntj) {
  return ti <= ntj ? ti * (ti + 1) / 2 : ntj * (ntj + 1) / 2 + (ti - ntj) * ntj;
}
SPK_HD inline int64_t schur_wide_before(int64_t ti, int64_t ntj) {
  const int64_t c = ntj / 2;  // row tiles whose 2 ti + 2 columns all lie below the cap
  return ti <= c ? ti * (ti + 1) : c * (c + 1) + (ti - c) * ntj;
}

// Per supernode with a Schur update: where its tasks {sn, ti, tj, d} go in
// the four Schur task arrays (64 x 64 tiles in plan order and longest first,
// 128 x 64 tiles in plan order and longest first).  nt = ceil(m / 64) row
// tiles, ntw = ceil(m / 128).
struct SchurEmit {         // mirrored on the device
  int64_t first[4];
  int32_t I, nt, ntw, ntj;
  int32_t d, nq, pad1, pad2;  // nq: k chunks of each task (panel and merged child)
};

struct LaunchRange {
  int kind;          // 0 diagonal factor, 7 panel TRSM, 9 both fused, 1 in-panel update,
                     // 2 Schur update,
                     // 3 forward diag, 4 forward off-diag, 5 backward off-diag, 6 backward diag,
                     // 10 forward / 11 backward partitioned-inverse solve, 12 its preparation
  int64_t first;     // first task in that kind's task array
  int64_t count;     // number of tasks (= grid size)
  int aux = 0;       // kinds 0, 7, 9, 3, 6: diagonal-block bucket (diag_bucket);
                     // kind 2: k-chunk split factor (0 or 1 = none)
  int level = 0;     // etree level of the launch (factor sequence)
  int alt = 0;       // kind 2: bit 0 = tasks in longest-first order, bit 1 = 128 x 64
                     // tiles (k_schur_wide), bit 2 = one wave per 64 x 64 tile (k_schur_w1)
  int64_t first_w = 0, count_w = 0;  // kind 2: the level's 128 x 64 tile tasks
  int nqmin = 0;     // kind 2: fewest k chunks of any task of the range
};

struct Plan {
  int32_t n = 0, nsuper = 0;
  int64_t arena_elems = 0;
  std::vector<SnInfo> sn;
  std::vector<int32_t> lindx0;     // 0-based structure rows (permuted numbering)
  std::vector<int32_t> level;      // etree height of every supernode
  int32_t nlevels = 0;
  // Schur scatter maps (per off-diagonal row b of supernode I):
  // target address of element (a,b) = tcol[upd+b] + rel[tmap[upd+b] + a].
  // tcol, tmap (upd_size entries each) and rel (rel_size entries) exist on
  // the device only, built from rel_groups.
  std::vector<RelGroup> rel_groups;
  int64_t upd_size = 0, rel_size = 0;
  // assembly: arena offset of every caller CSC entry (-1 = superseded duplicate)
  std::vector<int64_t> amap;
  std::vector<int32_t> perm0;      // perm0[new] = old, 0-based
  // task arrays
  std::vector<Task4> potrf_tasks;  // {sn, step, inverse-tile offset, 0}
  std::vector<Task4> trsm_tasks;   // {sn, step, row_tile, inverse-tile offset}
  int64_t inv_tiles = 0;           // 16x16 inverse diagonal tiles written by the factor
  std::vector<int64_t> inv_base;   // per supernode: first inverse tile (tile t of 16-column block t)
  std::vector<Task4> inpanel_tasks;// {sn, step, ti, tj}
  std::vector<Task4> fused_tasks;  // {sn, row tiles, row_tile, inverse-tile offset}: block factor + TRSM
  // Schur tasks {sn, ti, tj, d}: 64x64 output tiles; d = 0 plain, d > 0 the
  // update also carries merges[d-1] (a deferred child), d < 0 only output
  // columns b < -d (parent columns of a deferred supernode).  Four arrays,
  // written on the device from schur_emit: 64 x 64 tiles in plan order; the
  // same with every launch range's supernodes ordered by their k chunks,
  // longest first (stable: an alternative task order for the autotuner,
  // workgroups are dispatched in index order); 128 x 64 tiles {sn, 128-row
  // tile, 64-column tile, d} in plan order and longest first.
  int64_t n_schur = 0, n_schur_wide = 0;
  std::vector<SchurEmit> schur_emit;
  std::vector<MergeInfo> merges;
  std::vector<int32_t> mmap;
  std::vector<int32_t> deferred_into;  // per supernode: parent absorbing its update, -1 = none
  std::vector<LaunchRange> factor_seq;
  // solve
  std::vector<Task4> sdiag_tasks;  // {sn, 0, 0, 0}: diagonal-block solves
  std::vector<Task4> soffm_tasks;  // {sn, row_tile, 0, 0}: 64-row tiles of the off-diagonal panel
  std::vector<LaunchRange> fwd_seq, bwd_seq;  // kind 3/6 aux: diag_bucket of the level, 0 = w > kStep
  // Partitioned-inverse solve, available when every supernode is at most
  // kStep wide: G_I = [I; L_off] L_II^{-1} (r_I x w_I, the panel's shape and
  // offsets, in its own arena) is prepared once per factorization; then each
  // etree level costs one forward launch [X_I; dY] = G_I Y_I and one
  // backward launch X_I = G_I^T [X_I; -X_rows].  Tasks {sn, 64-row tile,
  // inverse-tile base, 0} over all rows of every panel; kinds 10/11 per
  // level, kind 12 (preparation) per diagonal-block bucket, aux = bucket.
  bool ginv_ok = false;
  std::vector<Task4> gtiles;
  std::vector<LaunchRange> gfwd_seq, gbwd_seq, gprep_seq;
  // dependency-driven solve (one persistent launch per sweep), available when
  // every supernode fits one diagonal block (w <= kStep).  Tasks
  // {kind, sn, row tile, first target}: forward D(s) = 0 / O(s,rt) = 1 in
  // postorder, backward R(s,rt) = 2 / B(s) = 3 in reverse postorder; the
  // targets of O(s,rt) / R(s,rt), df_tgt[d .. end), are the supernodes
  // owning the tile's rows.
  bool df_ok = false;  // the lists can be built (build_dataflow_lists)
  // algorithmic counts
  double schur_flops = 0, inpanel_flops = 0, diag_flops = 0;
  int64_t schur_entries = 0;       // lower-triangle entries scattered
  int max_w = 0;
  int64_t max_panel = 0;           // largest panel, r * w elements
};

// Build the plan from the symbolic analysis and the caller's CSC pattern
// (lower triangle, any base).  merge_updates: 0 no merged updates, 1 merge
// qualifying chain children on levels with enough tiles to fill the GPU,
// 2 merge every qualifying chain child.
void build_plan(const Symbolic& s, Int n, const Ptr* colptr, const Idx* rowind, int baseval,
                Plan& p, int merge_updates = 1);

// Task lists of the dependency-driven solve (switched off on the device):
// built on demand from the plan, not by build_plan.
struct DataflowLists {
  std::vector<Task4> fwd, bwd;
  std::vector<int32_t> fwd_end, bwd_end;
  std::vector<int32_t> tgt;          // shared by O and R tasks (same tiles)
  std::vector<int32_t> dcnt0;        // per supernode: O tasks that update its rows
  std::vector<int32_t> rcnt0;        // per supernode: its row tiles
};
DataflowLists build_dataflow_lists(const Plan& p);

// Checks the dependency-driven solve lists by running them in ticket order
// on one worker: every wait must already be satisfied when its task is
// reached (so with any number of workers taking tickets in order, the
// oldest unfinished task can always proceed), every counter must reach zero
// and every task's targets must be the owners of its rows.  Returns an
// empty string or the first violation.
std::string validate_solve_plan(const Plan& p);

}  // namespace spk
