// Host construction of the device plan (see plan.hpp).
#include "plan.hpp"

#include <algorithm>
#include <climits>
#include <stdexcept>
#include <exception>
#include <thread>

namespace spk {

namespace {
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
}  // namespace

void build_plan(const Symbolic& s, Int n, const Ptr* colptr, const Idx* rowind, int baseval,
                Plan& p, int merge_updates) {
  p = Plan();
  p.n = n;
  const Int ns = s.nsuper();
  p.nsuper = ns;
  p.sn.resize(ns);
  p.lindx0.resize(s.lindx.size());
  for (size_t i = 0; i < s.lindx.size(); ++i) p.lindx0[i] = s.lindx[i] - 1;
  p.perm0.resize(n);
  for (Int i = 0; i < n; ++i) p.perm0[i] = s.ord.perm[i] - 1;

  // panels, column-major, leading dimension r
  int64_t off = 0, upd = 0;
  for (Int I = 0; I < ns; ++I) {
    SnInfo& si = p.sn[I];
    si.fc = s.xsuper[I] - 1;
    si.w = s.xsuper[I + 1] - s.xsuper[I];
    si.lx = (int32_t)(s.xlindx[I] - 1);
    si.r = (int32_t)(s.xlindx[I + 1] - s.xlindx[I]);
    si.off = off;
    si.upd = upd;
    off += (int64_t)si.r * si.w;
    upd += si.r - si.w;
    p.max_w = std::max(p.max_w, (int)si.w);
    p.max_panel = std::max(p.max_panel, (int64_t)si.r * si.w);
  }
  p.arena_elems = off;

  // etree heights (children have smaller indices than parents: postorder)
  p.level.assign(ns, 0);
  for (Int I = 0; I < ns; ++I) {
    Int par = s.sparent[I];
    if (par > 0) p.level[par - 1] = std::max(p.level[par - 1], p.level[I] + 1);
  }
  p.nlevels = 0;
  for (Int I = 0; I < ns; ++I) p.nlevels = std::max(p.nlevels, p.level[I] + 1);

  // row -> supernode (0-based) and position maps for the relative indices
  std::vector<int32_t> row_sup(n);
  for (Int i = 0; i < n; ++i) row_sup[i] = s.supMembership[i] - 1;

  // Assembly map: caller entry (row,col) of the lower triangle goes to
  // (max,min) of the permuted indices; later duplicates win.  It runs on its
  // own thread beside the scatter maps and the merges (it writes only amap
  // and reads the structure).
  // Common case: every column's rows are strictly increasing and none lies
  // above the diagonal, so no two entries share a pair; each entry then
  // finds its row in its target supernode's structure by itself (the
  // diagonal block directly, the rows below by binary search), on several
  // threads over column ranges of equal entry counts.
  // Otherwise entries are bucketed by their permuted column (stable, in
  // entry order); each supernode marks the positions of its structure rows
  // once, so every entry finds its row in O(1).  Two entries land on the same
  // arena offset exactly when they share the permuted pair, i.e. the bucket
  // and the row, and the earlier one is dropped.
  bool amap_error = false;
  std::exception_ptr amap_exc;
  std::thread amap_thread([&]() {
    try {
    const Ptr nnz = colptr[n] - baseval;
    p.amap.resize(nnz);
    {
      const int nth = (int)std::max<unsigned>(1, std::min<unsigned>(8, std::thread::hardware_concurrency()));
      std::vector<Int> split(nth + 1, n);
      split[0] = 0;
      for (int t = 1; t < nth; ++t)
        split[t] = (Int)(std::lower_bound(colptr, colptr + n, (Ptr)(nnz * t / nth) + baseval) - colptr);
      std::vector<char> unclean(nth, 0), missing(nth, 0);
      auto work = [&](int t) {
        for (Int c = split[t]; c < split[t + 1]; ++c) {
          const int32_t pcol = s.ord.invp[c] - 1;
          Idx prev = -1;
          for (Ptr q = colptr[c] - baseval; q < colptr[c + 1] - baseval; ++q) {
            const Idx r = rowind[q] - baseval;
            if (r < c || r <= prev) {
              unclean[t] = 1;
              return;
            }
            prev = r;
            const int32_t prow = s.ord.invp[r] - 1;
            const int32_t pr = std::max(prow, pcol), pc = std::min(prow, pcol);
            const SnInfo& sj = p.sn[row_sup[pc]];
            int32_t pos = pr - sj.fc;
            if (pos >= sj.w) {
              const int32_t* b = &p.lindx0[sj.lx] + sj.w;
              const int32_t* e = &p.lindx0[sj.lx] + sj.r;
              const int32_t* f = std::lower_bound(b, e, pr);
              if (f == e || *f != pr) {
                missing[t] = 1;
                return;
              }
              pos = (int32_t)(f - &p.lindx0[sj.lx]);
            }
            p.amap[q] = sj.off + (int64_t)(pc - sj.fc) * sj.r + pos;
          }
        }
      };
      std::vector<std::thread> pool;
      for (int t = 1; t < nth; ++t) pool.emplace_back(work, t);
      work(0);
      for (auto& th : pool) th.join();
      bool any_unclean = false;
      for (int t = 0; t < nth; ++t) {
        any_unclean |= unclean[t] != 0;
        amap_error |= missing[t] != 0;
      }
      if (!any_unclean || amap_error) return;
    }
    std::vector<int32_t> apos(n, -1);
    {
      std::fill(p.amap.begin(), p.amap.end(), -1);
      std::vector<Ptr> bstart(n + 1, 0), order(nnz);
      std::vector<int32_t> trow_of(nnz);
      for (Int c = 0; c < n; ++c)
        for (Ptr q = colptr[c] - baseval; q < colptr[c + 1] - baseval; ++q) {
          const int32_t pr = s.ord.invp[rowind[q] - baseval] - 1, pc = s.ord.invp[c] - 1;
          trow_of[q] = std::max(pr, pc);
          ++bstart[std::min(pr, pc) + 1];
        }
      for (Int i = 0; i < n; ++i) bstart[i + 1] += bstart[i];
      {
        std::vector<Ptr> fill(bstart.begin(), bstart.end() - 1);
        for (Int c = 0; c < n; ++c)
          for (Ptr q = colptr[c] - baseval; q < colptr[c + 1] - baseval; ++q) {
            const int32_t pr = s.ord.invp[rowind[q] - baseval] - 1, pc = s.ord.invp[c] - 1;
            order[fill[std::min(pr, pc)]++] = q;
          }
      }
      std::vector<int32_t> seen(n, -1);
      std::vector<Ptr> last(n, -1);
      for (Int J = 0; J < ns; ++J) {
        const SnInfo& sj = p.sn[J];
        const int32_t* stJ = &p.lindx0[sj.lx];
        for (int32_t q = 0; q < sj.r; ++q) apos[stJ[q]] = q;
        for (int32_t tcol = sj.fc; tcol < sj.fc + sj.w; ++tcol)
          for (Ptr t = bstart[tcol]; t < bstart[tcol + 1]; ++t) {
            const Ptr q = order[t];
            const int32_t trow = trow_of[q];
            const int32_t pp = apos[trow];
            if (pp < 0) {
              amap_error = true;
              return;
            }
            p.amap[q] = sj.off + (int64_t)(tcol - sj.fc) * sj.r + pp;
            if (seen[trow] == tcol) p.amap[last[trow]] = -1;  // an earlier entry of the same pair
            seen[trow] = tcol;
            last[trow] = q;
          }
        for (int32_t q = 0; q < sj.r; ++q) apos[stJ[q]] = -1;
      }
    }
    } catch (...) {
      amap_exc = std::current_exception();
    }
  });
  struct Joiner {  // joins on every exit, also when the code below throws
    std::thread& t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } amap_joiner{amap_thread};

  // Schur scatter maps.  For off-diagonal row b of I with target supernode J
  // (the supernode of column lindx[w+b]) the rows a >= first row of that
  // target group are all in struct(J); their positions there (rel) are
  // found on the device from one RelGroup per group.
  p.upd_size = upd;
  p.rel_groups.clear();
  p.rel_size = 0;
  for (Int I = 0; I < ns; ++I) {
    const SnInfo& si = p.sn[I];
    const int32_t* st = &p.lindx0[si.lx];
    int32_t m = si.r - si.w;
    int32_t b = 0;
    while (b < m) {
      int32_t J = row_sup[st[si.w + b]];
      // group = consecutive off-diagonal rows falling in J's columns (the
      // columns of J are consecutive numbers: the group ends at the first
      // row past J's last column)
      int32_t gb = b;
      const int32_t jend = p.sn[J].fc + p.sn[J].w;
      b = (int32_t)(std::lower_bound(st + si.w + b, st + si.r, jend) - (st + si.w));
      p.rel_groups.push_back({p.rel_size, I, J, gb, b});
      p.rel_size += m - gb;
    }
  }

  // Deferred (merged) updates along supernode chains (see MergeInfo).  The
  // parent P of C is the supernode of C's first off-diagonal row; the rows of
  // struct(C) inside P's columns form a prefix of length n1, the remaining
  // rows R2 all lie in struct(P).  P absorbs at most one child, and a
  // supernode that absorbs a child is not itself deferred, so each merged
  // update has exactly two k segments.  A child qualifies when R2 covers most
  // of P's off-diagonal rows (the merged tiles then carry few zero rows).
  // Merging moves work from the child's level to the parent's and lengthens
  // the parent's tiles' k loop, which only pays where both levels have
  // enough tiles to keep every CU busy (3 workgroups per CU x 256 CUs); near
  // the root, where a level holds one or two supernodes, it would lengthen
  // the critical path instead.
  constexpr int32_t kMergeMinRows = 32;
  constexpr double kMergeMinFill = 0.85;
  const int64_t kMergeMinLevelTiles = merge_updates >= 2 ? 0 : 768;
  std::vector<int64_t> level_tiles(p.nlevels, 0);
  for (Int I = 0; I < ns; ++I) {
    const int64_t t = ceil_div(p.sn[I].r - p.sn[I].w, kTile);
    level_tiles[p.level[I]] += t * (t + 1) / 2;
  }
  std::vector<int32_t> pos(n, -1);
  p.deferred_into.assign(ns, -1);
  std::vector<int32_t> merge_of(ns, 0);  // per supernode: 1 + index into merges, 0 = none
  std::vector<int32_t> ncut(ns, 0);      // per deferred supernode: n1
  if (merge_updates > 0) {
    std::vector<int32_t> par(ns, -1), n1(ns, 0);
    std::vector<std::vector<Int>> kids(ns);
    for (Int I = 0; I < ns; ++I) {
      const SnInfo& si = p.sn[I];
      const int32_t m = si.r - si.w;
      if (m <= 0) continue;
      const int32_t* st = &p.lindx0[si.lx + si.w];
      const Int P = row_sup[st[0]];
      const int32_t end = p.sn[P].fc + p.sn[P].w;
      int32_t k = 0;
      while (k < m && st[k] < end) ++k;
      par[I] = P;
      n1[I] = k;
      kids[P].push_back(I);
    }
    std::vector<char> is_merger(ns, 0);
    std::vector<int32_t> map;
    for (Int P = 0; P < ns; ++P) {  // children precede parents
      const SnInfo& sp = p.sn[P];
      const int32_t mP = sp.r - sp.w;
      if (mP <= 0 || level_tiles[p.level[P]] < kMergeMinLevelTiles) continue;
      Int best = -1;
      int32_t bv = 0;
      for (Int C : kids[P]) {
        if (is_merger[C] || level_tiles[p.level[C]] < kMergeMinLevelTiles) continue;
        const int32_t r2 = (p.sn[C].r - p.sn[C].w) - n1[C];
        if (r2 >= kMergeMinRows && (double)r2 >= kMergeMinFill * mP && r2 > bv) {
          best = C;
          bv = r2;
        }
      }
      if (best < 0) continue;
      const SnInfo& sc = p.sn[best];
      const int32_t* stC = &p.lindx0[sc.lx];
      for (int32_t q = 0; q < sc.r; ++q) pos[stC[q]] = q;
      map.assign(mP, -1);
      int32_t found = 0;
      for (int32_t a = 0; a < mP; ++a) {
        const int32_t q = pos[p.lindx0[sp.lx + sp.w + a]];
        if (q >= sc.w) {
          map[a] = q;
          ++found;
        }
      }
      for (int32_t q = 0; q < sc.r; ++q) pos[stC[q]] = -1;
      if (found != bv) continue;  // R2 not inside struct(P): keep the plain update
      MergeInfo mi{};
      mi.off = sc.off;
      mi.moff = (int64_t)p.mmap.size();
      mi.ld = sc.r;
      mi.w = sc.w;
      mi.child = best;
      p.mmap.insert(p.mmap.end(), map.begin(), map.end());
      is_merger[P] = 1;
      p.deferred_into[best] = P;
      ncut[best] = n1[best];
      merge_of[P] = (int32_t)p.merges.size() + 1;
      p.merges.push_back(mi);
    }
  }

  amap_thread.join();
  if (amap_exc) std::rethrow_exception(amap_exc);
  if (amap_error) throw std::logic_error("matrix entry outside the symbolic structure");

  // inverse diagonal tiles: ceil(w/16) per supernode, contiguous, so block t
  // of step st is tile inv_base + 10 st + t
  p.inv_base.assign(ns, 0);
  for (Int I = 0; I < ns; ++I) {
    p.inv_base[I] = p.inv_tiles;
    p.inv_tiles += ceil_div(p.sn[I].w, kSub);
  }

  // ---- factorization task lists, level by level ----
  std::vector<std::vector<Int>> by_level(p.nlevels);
  for (Int I = 0; I < ns; ++I) by_level[p.level[I]].push_back(I);
  for (int L = 0; L < p.nlevels; ++L) {
    const auto& sns = by_level[L];
    const size_t level_first = p.factor_seq.size();
    int smax = 0;
    for (Int I : sns) smax = std::max<int>(smax, (int)ceil_div(p.sn[I].w, kStep));
    // Levels of single-step supernodes with at most kFuseMaxTasks row
    // tiles in all: one fused block-factor + TRSM launch (every row-tile
    // workgroup factors the block again) instead of two dependent launches.
    constexpr int64_t kFuseMaxTasks = 256;
    int64_t fuse_tasks = 0;
    if (smax == 1)
      for (Int I : sns) fuse_tasks += std::max<int64_t>(1, ceil_div(p.sn[I].r - p.sn[I].w, kTrsmRows));
    if (smax == 1 && fuse_tasks <= kFuseMaxTasks) {
      LaunchRange fr{9, (int64_t)p.fused_tasks.size(), 0};
      int nbmax = 1;
      for (Int I : sns) {
        const SnInfo& si = p.sn[I];
        const int bs = si.w;
        nbmax = std::max(nbmax, (int)ceil_div(bs, kSub));
        const int64_t rows_below = si.r - bs;
        const int32_t inv_off = (int32_t)p.inv_base[I];
        const int64_t nt = std::max<int64_t>(1, ceil_div(rows_below, kTrsmRows));
        for (int64_t rt = 0; rt < nt; ++rt)
          p.fused_tasks.push_back({I, (int32_t)nt, (int32_t)rt, inv_off});
        p.diag_flops += (double)bs * bs * bs / 3.0 + (double)bs * bs * rows_below;
      }
      fr.aux = diag_bucket(nbmax);
      fr.count = (int64_t)p.fused_tasks.size() - fr.first;
      p.factor_seq.push_back(fr);
      smax = 0;  // nothing left for the per-step launches below
    }
    for (int st = 0; st < smax; ++st) {
      LaunchRange pr{0, (int64_t)p.potrf_tasks.size(), 0};
      LaunchRange tr{7, (int64_t)p.trsm_tasks.size(), 0};
      LaunchRange ur{1, (int64_t)p.inpanel_tasks.size(), 0};
      int nbmax = 1;
      for (Int I : sns) {
        const SnInfo& si = p.sn[I];
        int nsteps = (int)ceil_div(si.w, kStep);
        if (st >= nsteps) continue;
        int c0 = st * kStep, bs = std::min(kStep, si.w - c0);
        int nb = (int)ceil_div(bs, kSub);
        nbmax = std::max(nbmax, nb);
        int64_t rows_below = si.r - c0 - bs;
        int32_t inv_off = (int32_t)(p.inv_base[I] + (int64_t)st * (kStep / kSub));
        p.potrf_tasks.push_back({I, st, inv_off, 0});
        for (int64_t rt = 0; rt < ceil_div(rows_below, kTrsmRowsWide); ++rt)
          p.trsm_tasks.push_back({I, st, (int32_t)rt, inv_off});
        p.diag_flops += (double)bs * bs * bs / 3.0 + (double)bs * bs * rows_below;
        if (st < nsteps - 1) {
          int rb = c0 + bs;  // trailing rows/cols start
          int64_t nrow = si.r - rb, ncol = si.w - rb;
          int64_t ntr = ceil_div(nrow, kTile), ntc = ceil_div(ncol, kTile);
          for (int64_t ti = 0; ti < ntr; ++ti)
            for (int64_t tj = 0; tj < ntc && tj <= ti; ++tj)
              p.inpanel_tasks.push_back({I, st, (int32_t)ti, (int32_t)tj});
          // flops of the lower-trapezoid update with k = bs
          p.inpanel_flops += (double)bs * ((double)ncol * (ncol + 1) + 2.0 * (nrow - ncol) * ncol);
        }
      }
      pr.aux = tr.aux = diag_bucket(nbmax);
      pr.count = (int64_t)p.potrf_tasks.size() - pr.first;
      tr.count = (int64_t)p.trsm_tasks.size() - tr.first;
      ur.count = (int64_t)p.inpanel_tasks.size() - ur.first;
      if (pr.count) p.factor_seq.push_back(pr);
      if (tr.count) p.factor_seq.push_back(tr);
      if (ur.count) p.factor_seq.push_back(ur);
    }
    // Schur tasks: counted here, written on the device (schur_emit)
    LaunchRange sr{2, p.n_schur, 0};
    sr.first_w = p.n_schur_wide;
    const size_t emit_first = p.schur_emit.size();
    int nqmin = INT_MAX, nqmax = 0;
    for (Int I : sns) {
      const SnInfo& si = p.sn[I];
      int64_t m = si.r - si.w;
      if (m <= 0) continue;
      SchurEmit e{};
      e.I = I;
      e.nt = (int32_t)ceil_div(m, kTile);
      e.ntw = (int32_t)ceil_div(m, 2 * kTile);
      if (p.deferred_into[I] >= 0) {
        // only the output columns in the parent's columns; the rest is
        // computed by the parent's merged update
        const int32_t n1 = ncut[I];
        e.ntj = (int32_t)ceil_div(n1, kTile);
        e.d = -n1;
        p.schur_entries += (int64_t)n1 * m - (int64_t)n1 * (n1 - 1) / 2;
      } else {
        e.ntj = e.nt;
        e.d = merge_of[I];
        p.schur_entries += m * (m + 1) / 2;
      }
      e.first[0] = p.n_schur;
      e.first[2] = p.n_schur_wide;
      p.n_schur += schur_tiles_before(e.nt, e.ntj);
      p.n_schur_wide += schur_wide_before(e.ntw, e.ntj);
      int nq = (int)ceil_div(si.w, kSchurKC);
      if (e.d > 0) nq += (int)ceil_div(p.merges[e.d - 1].w, kSchurKC);
      nqmin = std::min(nqmin, nq);
      nqmax = std::max(nqmax, nq);
      e.nq = nq;
      p.schur_emit.push_back(e);
      // algorithmic flops stay those of the supernode's whole update, wherever
      // its deferred part is computed
      p.schur_flops += (double)si.w * (double)m * (double)(m + 1);
    }
    sr.count = p.n_schur - sr.first;
    sr.count_w = p.n_schur_wide - sr.first_w;
    sr.nqmin = sr.count ? nqmin : 0;
    // longest-first order: the level's supernodes by k chunks, longest first,
    // stable (a counting sort on the chunk count, a small integer), each
    // supernode's tiles in their plan order
    {
      std::vector<int64_t> cnt(nqmax + 2, 0);
      for (size_t q = emit_first; q < p.schur_emit.size(); ++q) ++cnt[nqmax - p.schur_emit[q].nq + 1];
      for (int i = 1; i <= nqmax + 1; ++i) cnt[i] += cnt[i - 1];
      std::vector<size_t> order(p.schur_emit.size() - emit_first);
      for (size_t q = emit_first; q < p.schur_emit.size(); ++q)
        order[cnt[nqmax - p.schur_emit[q].nq]++] = q;
      int64_t f1 = sr.first, f3 = sr.first_w;
      for (size_t q : order) {
        SchurEmit& e = p.schur_emit[q];
        e.first[1] = f1;
        e.first[3] = f3;
        f1 += schur_tiles_before(e.nt, e.ntj);
        f3 += schur_wide_before(e.ntw, e.ntj);
      }
    }
    if (sr.count && sr.count < kSplitTarget) {
      // few output tiles: each task's k chunks are split over S workgroups
      // (partial sums meet in the atomic scatter), S <= the fewest chunks of
      // any task of the level so that no part is empty
      sr.aux = (int)std::max<int64_t>(1, std::min<int64_t>(nqmin, kSplitTarget / sr.count));
    }
    if (sr.count) p.factor_seq.push_back(sr);
    for (size_t q = level_first; q < p.factor_seq.size(); ++q) p.factor_seq[q].level = L;
  }

  // ---- solve task lists ----
  // Per level: one diagonal task per supernode ({sn, inverse-tile base}) and
  // one off-diagonal task per 64-row tile of the panel below the diagonal.
  std::vector<int64_t> dfirst(p.nlevels), mfirst(p.nlevels), mcount(p.nlevels);
  for (int L = 0; L < p.nlevels; ++L) {
    const auto& sns = by_level[L];
    dfirst[L] = (int64_t)p.sdiag_tasks.size();
    for (Int I : sns) p.sdiag_tasks.push_back({I, (int32_t)p.inv_base[I], 0, 0});
    mfirst[L] = (int64_t)p.soffm_tasks.size();
    for (Int I : sns) {
      int64_t m = p.sn[I].r - p.sn[I].w;
      for (int64_t rt = 0; rt < ceil_div(m, kTile); ++rt) p.soffm_tasks.push_back({I, (int32_t)rt, 0, 0});
    }
    mcount[L] = (int64_t)p.soffm_tasks.size() - mfirst[L];
  }
  // partitioned-inverse solve: 64-row tiles over all rows of every panel,
  // grouped by level (solve sweeps) and by diagonal-block bucket (the
  // preparation, one launch per bucket over all supernodes)
  p.ginv_ok = p.max_w <= kStep;
  if (p.ginv_ok) {
    std::vector<int> lbucket(p.nlevels, 1);
    for (int L = 0; L < p.nlevels; ++L)
      for (Int I : by_level[L]) lbucket[L] = std::max(lbucket[L], diag_bucket((int)ceil_div(p.sn[I].w, kSub)));
    std::vector<int64_t> gfirst(p.nlevels), gcount(p.nlevels);
    for (int L = 0; L < p.nlevels; ++L) {
      gfirst[L] = (int64_t)p.gtiles.size();
      for (Int I : by_level[L])
        for (int64_t rt = 0; rt < ceil_div(p.sn[I].r, kTile); ++rt)
          p.gtiles.push_back({I, (int32_t)rt, (int32_t)p.inv_base[I], 0});
      gcount[L] = (int64_t)p.gtiles.size() - gfirst[L];
    }
    for (int L = 0; L < p.nlevels; ++L) {
      LaunchRange f{10, gfirst[L], gcount[L], lbucket[L], L};
      if (gcount[L]) p.gfwd_seq.push_back(f);
    }
    for (int L = p.nlevels - 1; L >= 0; --L) {
      LaunchRange b{11, gfirst[L], gcount[L], lbucket[L], L};
      if (gcount[L]) p.gbwd_seq.push_back(b);
    }
    // preparation: the same tiles, regrouped by the bucket of their own
    // supernode (appended after the level-ordered tiles)
    for (int bk : {1, 2, 4, 6, 10}) {
      LaunchRange r{12, (int64_t)p.gtiles.size(), 0, bk, -1};
      for (Int I = 0; I < ns; ++I) {
        if (diag_bucket((int)ceil_div(p.sn[I].w, kSub)) != bk) continue;
        for (int64_t rt = 0; rt < ceil_div(p.sn[I].r, kTile); ++rt)
          p.gtiles.push_back({I, (int32_t)rt, (int32_t)p.inv_base[I], 0});
      }
      r.count = (int64_t)p.gtiles.size() - r.first;
      if (r.count) p.gprep_seq.push_back(r);
    }
  }

  // the dependency-driven solve's lists are built on demand (build_dataflow_lists)
  p.df_ok = p.max_w <= kStep;

  // forward: levels bottom-up, diagonal solves then off-diagonal updates.
  std::vector<int> dbucket(p.nlevels, 1);
  for (int L = 0; L < p.nlevels; ++L) {
    int wmax = 1;
    for (Int I : by_level[L]) wmax = std::max<int>(wmax, p.sn[I].w);
    dbucket[L] = wmax > kStep ? 0 : diag_bucket((int)ceil_div(wmax, kSub));
  }
  for (int L = 0; L < p.nlevels; ++L) {
    LaunchRange d{3, dfirst[L], (int64_t)by_level[L].size()};
    d.aux = dbucket[L];
    d.level = L;
    p.fwd_seq.push_back(d);
    if (mcount[L]) p.fwd_seq.push_back({4, mfirst[L], mcount[L], 0, L});
  }
  // backward: levels top-down, off-diagonal gathers then diagonal solves.
  for (int L = p.nlevels - 1; L >= 0; --L) {
    if (mcount[L]) p.bwd_seq.push_back({5, mfirst[L], mcount[L], 0, L});
    LaunchRange d{6, dfirst[L], (int64_t)by_level[L].size()};
    d.aux = dbucket[L];
    d.level = L;
    p.bwd_seq.push_back(d);
  }
}


DataflowLists build_dataflow_lists(const Plan& p) {
  DataflowLists L;
  if (!p.df_ok) return L;
  const Int ns = p.nsuper;
  std::vector<int32_t> row_sup(p.n);
  for (Int I = 0; I < ns; ++I)
    for (int32_t c = p.sn[I].fc; c < p.sn[I].fc + p.sn[I].w; ++c) row_sup[c] = I;

    L.dcnt0.assign(ns, 0);
    L.rcnt0.assign(ns, 0);
    std::vector<int32_t> last_seen(ns, -1);
    std::vector<int64_t> dtask(ns, 0);  // forward D task of every supernode
    int32_t stamp = 0;
    for (Int I = 0; I < ns; ++I) {
      const SnInfo& si = p.sn[I];
      const int32_t m = si.r - si.w;
      const int32_t nt = (int32_t)ceil_div(m, kTile);
      L.rcnt0[I] = nt;
      dtask[I] = (int64_t)L.fwd.size();
      L.fwd.push_back({0, I, 0, 0});
      L.fwd_end.push_back(0);
      for (int32_t rt = 0; rt < nt; ++rt) {
        const int32_t first = (int32_t)L.tgt.size();
        ++stamp;
        for (int32_t a = rt * kTile; a < std::min<int32_t>(m, (rt + 1) * kTile); ++a) {
          const int32_t T = row_sup[p.lindx0[si.lx + si.w + a]];
          if (last_seen[T] != stamp) {
            last_seen[T] = stamp;
            L.tgt.push_back(T);
            ++L.dcnt0[T];
          }
        }
        L.fwd.push_back({1, I, rt, first});
        L.fwd_end.push_back((int32_t)L.tgt.size());
      }
    }
    // backward: the same tiles in reverse postorder, waiting for the owners
    // of their rows
    for (Int I = ns - 1; I >= 0; --I) {
      for (int32_t rt = 0; rt < L.rcnt0[I]; ++rt) {
        const int64_t q = dtask[I] + 1 + rt;
        L.bwd.push_back({2, I, rt, L.fwd[q].d});
        L.bwd_end.push_back(L.fwd_end[q]);
      }
      L.bwd.push_back({3, I, 0, 0});
      L.bwd_end.push_back(0);
    }
  return L;
}

std::string validate_solve_plan(const Plan& p) {
  if (!p.df_ok) return std::string();
  const DataflowLists L = build_dataflow_lists(p);
  const Int ns = p.nsuper;
  std::vector<int32_t> dcnt = L.dcnt0, rcnt = L.rcnt0;
  std::vector<char> dflag(ns, 0), bflag(ns, 0);
  std::vector<int32_t> row_sup(p.n, -1);
  for (Int I = 0; I < ns; ++I)
    for (int32_t c = 0; c < p.sn[I].w; ++c) row_sup[p.sn[I].fc + c] = I;
  auto owners_ok = [&](const Task4& t, int32_t end) -> bool {
    const SnInfo& si = p.sn[t.b];
    const int32_t m = si.r - si.w;
    std::vector<int32_t> want;
    for (int32_t a = t.c * kTile; a < std::min<int32_t>(m, (t.c + 1) * kTile); ++a) {
      const int32_t T = row_sup[p.lindx0[si.lx + si.w + a]];
      if (std::find(want.begin(), want.end(), T) == want.end()) want.push_back(T);
    }
    std::vector<int32_t> got(L.tgt.begin() + t.d, L.tgt.begin() + end);
    std::sort(want.begin(), want.end());
    std::sort(got.begin(), got.end());
    return want == got;
  };
  for (size_t q = 0; q < L.fwd.size(); ++q) {
    const Task4& t = L.fwd[q];
    if (t.a == 0) {
      if (dcnt[t.b] != 0) return "forward D(" + std::to_string(t.b) + ") reached with pending updates";
      dflag[t.b] = 1;
    } else if (t.a == 1) {
      if (!dflag[t.b]) return "forward O(" + std::to_string(t.b) + ") before its diagonal solve";
      if (!owners_ok(t, L.fwd_end[q])) return "forward O(" + std::to_string(t.b) + ") targets";
      for (int32_t k = t.d; k < L.fwd_end[q]; ++k) --dcnt[L.tgt[k]];
    } else {
      return "bad forward task kind";
    }
  }
  for (Int I = 0; I < ns; ++I)
    if (dcnt[I] != 0 || !dflag[I]) return "forward counters not drained";
  for (size_t q = 0; q < L.bwd.size(); ++q) {
    const Task4& t = L.bwd[q];
    if (t.a == 2) {
      for (int32_t k = t.d; k < L.bwd_end[q]; ++k)
        if (!bflag[L.tgt[k]]) return "backward R(" + std::to_string(t.b) + ") before an owner is solved";
      if (!owners_ok(t, L.bwd_end[q])) return "backward R(" + std::to_string(t.b) + ") targets";
      --rcnt[t.b];
    } else if (t.a == 3) {
      if (rcnt[t.b] != 0) return "backward B(" + std::to_string(t.b) + ") reached with pending tiles";
      bflag[t.b] = 1;
    } else {
      return "bad backward task kind";
    }
  }
  for (Int I = 0; I < ns; ++I)
    if (rcnt[I] != 0 || !bflag[I]) return "backward counters not drained";
  return std::string();
}

}  // namespace spk
