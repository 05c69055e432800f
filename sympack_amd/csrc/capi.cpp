// C ABI over the engine (declarations and reference citations in
// include/sympack_amd.h).  Exceptions never cross this boundary: they become
// negative status codes with the message kept for spk_last_error().
#include "../../include/sympack_amd.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

#include "analysis/analysis.hpp"
#include "engine/engine.hpp"

struct spk_matrix_s {
  spk::Engine engine;
  spk::EngineOptions opt;
  const int* user_perm = nullptr;  // borrowed, as symPACKOptions.perm
  spk::CellMap cells;
  bool cells_built = false;
};

namespace {
thread_local std::string g_last_error;
int g_device = 0;
bool g_initialized = false;

int fail(const std::string& msg, int code = -1) {
  g_last_error = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::logic_error& e) {
    return fail(std::string("logic_error: ") + e.what(), -2);
  } catch (const std::bad_alloc& e) {
    return fail(std::string("bad_alloc: ") + e.what(), -3);
  } catch (const std::exception& e) {
    return fail(std::string("runtime_error: ") + e.what(), -1);
  } catch (...) {
    return fail("unknown error", -1);
  }
}

spk::EngineOptions to_engine(const spk_options& o, int n_hint) {
  (void)n_hint;
  spk::EngineOptions e;
  e.ordering = o.ordering;
  e.ldl = o.decomposition == SPK_FACT_LDL;
  e.relax.nrelax0 = o.nrelax0;
  e.relax.nrelax1 = o.nrelax1;
  e.relax.nrelax2 = o.nrelax2;
  e.relax.maxSize = o.maxsize;
  // RelaxationParameters::SetMaxSize clamps the nrelax values (Types.hpp:62-69)
  if (o.maxsize > 0) {
    e.relax.nrelax0 = std::min(e.relax.nrelax0, o.maxsize);
    e.relax.nrelax1 = std::min(e.relax.nrelax1, o.maxsize);
    e.relax.nrelax2 = std::min(e.relax.nrelax2, o.maxsize);
  }
  e.relax.zrelax0 = o.zrelax0;
  e.relax.zrelax1 = o.zrelax1;
  e.relax.zrelax2 = o.zrelax2;
  e.precision = o.precision;
  e.ir_max_iter = o.ir_max_iter;
  e.ir_tol = o.ir_tol;
  e.verbose = o.verbose;
  e.profile = o.profile;
  e.device = o.device;
  e.merge_updates = o.merge_updates;
  e.solve_dataflow = o.solve_dataflow;
  // the reordering codes ordsup takes: 10 * ordflag + altflag
  // (symPACKMatrix2D.hpp:3090-3111); anything else is an argument error
  if (o.refine != 0 && o.refine != 10 && o.refine != 11 && o.refine != 20 && o.refine != 21 &&
      o.refine != 30 && o.refine != 31 && o.refine != spk::kRefineTspb && o.refine != spk::kRefineTspbPlain)
    throw std::logic_error("refine must be 0, 10, 11, 20, 21, 30, 31, 40 or 41, got " + std::to_string(o.refine));
  if (o.precision != SPK_PREC_FP64 && o.precision != SPK_PREC_FP32_IR)
    throw std::logic_error("precision must be SPK_PREC_FP64 or SPK_PREC_FP32_IR");
  e.refine = o.refine;
  e.solve_inverse = o.solve_inverse;
  e.autotune = o.autotune;
  if (o.schur_kernel < 0 || o.schur_kernel > 3)
    throw std::logic_error("schur_kernel must be 0, 1, 2 or 3, got " + std::to_string(o.schur_kernel));
  e.schur_kernel = o.schur_kernel;
  return e;
}
}  // namespace

extern "C" {

int symPACK_Init(int* argc, char*** argv) {
  (void)argc;
  (void)argv;
  g_initialized = true;
  return 0;
}

int symPACK_Finalize(void) {
  g_initialized = false;
  return 0;
}

int symPACK_Rank(int* rank) {
  if (!rank) return fail("null rank pointer");
  *rank = 0;
  return 0;
}

int symPACK_device_setup(int device) {
  int cnt = 0;
  hipError_t e = hipGetDeviceCount(&cnt);
  if (e != hipSuccess) return fail("hipGetDeviceCount: " + spk::hip_error_string(e));
  if (device < 0 || device >= cnt) return fail("device ordinal out of range");
  g_device = device;
  e = hipSetDevice(device);
  if (e != hipSuccess) return fail("hipSetDevice: " + spk::hip_error_string(e));
  return 0;
}

void spk_default_options(spk_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->ordering = SPK_ORDER_MMD;
  o->decomposition = SPK_FACT_LL;
  o->nrelax0 = 8;
  o->nrelax1 = 32;
  o->nrelax2 = 64;
  o->maxsize = 0;
  o->zrelax0 = 0.8;
  o->zrelax1 = 0.1;
  o->zrelax2 = 0.05;
  o->precision = SPK_PREC_FP64;
  o->ir_max_iter = 30;
  o->ir_tol = 1e-12;
  o->perm = nullptr;
  o->verbose = 0;
  o->profile = 0;
  o->device = g_device;
  o->merge_updates = 1;
  o->solve_dataflow = 0;
  o->refine = 0;
  o->solve_inverse = 1;
  o->autotune = 0;
  o->schur_kernel = 0;
}

int spk_options_from_args(spk_options* o, int argc, const char* const* argv) {
  if (!o) return fail("null options");
  return guarded([&]() {
    // OptionsCreate (Environment_impl.hpp:24-42): a "-key" starts a list of values.
    for (int i = 0; i < argc; ++i) {
      std::string k = argv[i];
      if (k.empty() || k[0] != '-') continue;
      auto val = [&](int off) -> std::string {
        if (i + off < argc && argv[i + off][0] != '-') return argv[i + off];
        if (i + off < argc) {  // negative numbers are values too
          const char* v = argv[i + off];
          if (v[1] >= '0' && v[1] <= '9') return v;
        }
        return std::string();
      };
      if (k == "-ordering") {
        std::string v = val(1);
        if (v == "MMD") o->ordering = SPK_ORDER_MMD;
        else if (v == "RCM") o->ordering = SPK_ORDER_RCM;
        else if (v == "AMD") o->ordering = SPK_ORDER_AMD;
        else if (v == "NATURAL") o->ordering = SPK_ORDER_NATURAL;
        else if (v == "NDBOX") o->ordering = SPK_ORDER_NDBOX;
        else if (v == "NDGRID") o->ordering = SPK_ORDER_NDGRID;
        else if (v == "USER") o->ordering = SPK_ORDER_USER;
        else
          throw std::logic_error(
              "This ordering method is not supported by symPACK. Valid options "
              "are:NATURAL MMD AMD RCM NDBOX NDGRID USER \n");
      } else if (k == "-fact") {
        std::string v = val(1);
        if (v == "LL") o->decomposition = SPK_FACT_LL;
        else if (v == "LDL") o->decomposition = SPK_FACT_LDL;
      } else if (k == "-relax") {
        std::string a = val(1), b = val(2), c = val(3);
        if (!a.empty() && !b.empty() && !c.empty()) {
          o->nrelax0 = std::atoi(a.c_str());
          o->nrelax1 = std::atoi(b.c_str());
          o->nrelax2 = std::atoi(c.c_str());
        } else {
          o->nrelax0 = 0;  // disable relaxed supernodes (utils.hpp:354-357)
        }
      } else if (k == "-b") {
        std::string v = val(1);
        o->maxsize = (v == "inf" || v.empty()) ? 0 : std::atoi(v.c_str());
      } else if (k == "-v") {
        o->verbose = std::atoi(val(1).c_str());
      } else if (k == "-precision") {
        std::string v = val(1);
        if (v == "FP64" || v == "fp64")
          o->precision = SPK_PREC_FP64;
        else if (v == "FP32" || v == "fp32" || v == "FP32IR" || v == "FP32_IR" || v == "fp32_ir")
          o->precision = SPK_PREC_FP32_IR;
        else
          throw std::logic_error("-precision must be FP64 or FP32_IR (FP32, FP32IR), got '" + v + "'");
      } else if (k == "-ir_tol") {
        o->ir_tol = std::atof(val(1).c_str());
      } else if (k == "-ir_iter") {
        o->ir_max_iter = std::atoi(val(1).c_str());
      } else if (k == "-merge") {
        o->merge_updates = std::atoi(val(1).c_str());
      } else if (k == "-refine") {
        // symPACKMatrix2D.hpp:3090-3263 (utils.hpp:331-333): SET = SET31,
        // TSPB = tour with skeleton compression, other TSPB* names = tour
        // without it; other names leave the ordering as it is
        std::string v = val(1);
        if (v == "SET" || v == "SET31") o->refine = 31;
        else if (v == "SET10") o->refine = 10;
        else if (v == "SET11") o->refine = 11;
        else if (v == "SET20") o->refine = 20;
        else if (v == "SET21") o->refine = 21;
        else if (v == "SET30") o->refine = 30;
        else if (v == "TSPB") o->refine = spk::kRefineTspb;
        else if (v.compare(0, 4, "TSPB") == 0) o->refine = spk::kRefineTspbPlain;
        else o->refine = 0;
      } else if (k == "-solve_inverse") {
        o->solve_inverse = std::atoi(val(1).c_str()) != 0;
      } else if (k == "-autotune") {
        o->autotune = std::atoi(val(1).c_str()) != 0;
      } else if (k == "-schur_kernel") {
        o->schur_kernel = std::atoi(val(1).c_str());
      } else if (k == "-dataflow") {
        o->solve_dataflow = std::atoi(val(1).c_str()) != 0;
      }
    }
    return 0;
  });
}

int spk_create(spk_handle* h) {
  if (!h) return fail("null handle pointer");
  return guarded([&]() {
    *h = new spk_matrix_s();
    spk_options o;
    spk_default_options(&o);
    (*h)->opt = to_engine(o, 0);
    return 0;
  });
}

int spk_destroy(spk_handle h) {
  delete h;
  return 0;
}

int spk_init(spk_handle h, const spk_options* opt) {
  if (!h || !opt) return fail("null argument");
  return guarded([&]() {
    h->opt = to_engine(*opt, 0);
    h->user_perm = opt->perm;
    h->engine.init(h->opt);
    return 0;
  });
}

int spk_symbolic_factorization(spk_handle h, int32_t n, const int64_t* colptr,
                               const int32_t* rowind, int baseval) {
  if (!h || !colptr || (!rowind && n > 0)) return fail("null argument");
  if (n <= 0) return fail("matrix dimension must be positive");
  if (baseval != 0 && baseval != 1) return fail("baseval must be 0 or 1");
  return guarded([&]() {
    h->cells_built = false;
    std::vector<spk::Int> up;
    if (h->opt.ordering == SPK_ORDER_USER && h->user_perm) up.assign(h->user_perm, h->user_perm + n);
    h->engine.set_user_perm(up);
    h->engine.symbolic(n, colptr, rowind, baseval);
    return 0;
  });
}

int spk_distribute_matrix(spk_handle h, const double* nzval) {
  if (!h || !nzval) return fail("null argument");
  return guarded([&]() {
    h->engine.distribute(nzval, false);
    return 0;
  });
}

int spk_distribute_matrix_device(spk_handle h, const double* d_nzval) {
  if (!h || !d_nzval) return fail("null argument");
  return guarded([&]() {
    h->engine.distribute(d_nzval, true);
    return 0;
  });
}

int spk_factorize(spk_handle h) {
  if (!h) return fail("null handle");
  return guarded([&]() {
    int info = h->engine.factorize();
    if (info > 0) g_last_error = "Matrix is not HPD.";
    return info;
  });
}

int spk_solve(spk_handle h, const double* rhs, int nrhs, int64_t rhs_size) {
  (void)rhs_size;
  if (!h || !rhs) return fail("null argument");
  if (nrhs <= 0) return fail("nrhs must be positive");
  return guarded([&]() {
    h->engine.solve(rhs, nrhs, false);
    return 0;
  });
}

int spk_solve_device(spk_handle h, const double* d_rhs, int nrhs) {
  if (!h || !d_rhs) return fail("null argument");
  if (nrhs <= 0) return fail("nrhs must be positive");
  return guarded([&]() {
    h->engine.solve(d_rhs, nrhs, true);
    return 0;
  });
}

int spk_solve_device_noalloc(spk_handle h, const double* d_rhs, double* d_x, int nrhs) {
  if (!h || !d_rhs || !d_x) return fail("null argument");
  if (nrhs <= 0) return fail("nrhs must be positive");
  return guarded([&]() {
    h->engine.solve_noalloc(d_rhs, d_x, nrhs);
    return 0;
  });
}

int spk_get_solution(spk_handle h, double* B, int nrhs) {
  if (!h || !B) return fail("null argument");
  return guarded([&]() {
    h->engine.get_solution(B, nrhs, false);
    return 0;
  });
}

int spk_get_solution_device(spk_handle h, double* d_B, int nrhs) {
  if (!h || !d_B) return fail("null argument");
  return guarded([&]() {
    h->engine.get_solution(d_B, nrhs, true);
    return 0;
  });
}

int spk_get_stats(spk_handle h, spk_stats* st) {
  if (!h || !st) return fail("null argument");
  return guarded([&]() {
    const spk::Plan& p = h->engine.plan();
    const spk::Symbolic& s = h->engine.sym();
    std::memset(st, 0, sizeof(*st));
    st->n = p.n;
    st->nsuper = p.nsuper;
    st->nlevels = p.nlevels;
    st->max_width = p.max_w;
    st->flops = s.flops;
    st->nnz_l = s.nnzL;
    st->factor_entries = p.arena_elems;
    st->schur_flops = p.schur_flops;
    st->inpanel_flops = p.inpanel_flops;
    st->diag_flops = p.diag_flops;
    st->schur_entries = p.schur_entries;
    const auto& ks = h->engine.kind_stats;
    st->launches_factor = ks[0].launches + ks[1].launches + ks[2].launches + ks[7].launches + ks[9].launches;
    st->launches_solve = ks[3].launches + ks[4].launches + ks[5].launches + ks[6].launches +
                         ks[10].launches + ks[11].launches + ks[12].launches;
    st->factor_ms = h->engine.last_factor_ms;
    st->solve_ms = h->engine.last_solve_ms;
    st->ir_iterations = h->engine.last_report.iterations;
    st->backward_error = h->engine.last_report.backward_error;
    st->fp32_shift = h->engine.fp32_shift_;
    st->ir_converged = h->engine.last_report.converged;
    return 0;
  });
}

int spk_get_kernel_stats(spk_handle h, int kind, int64_t* launches, double* ms) {
  if (!h || !launches || !ms) return fail("null argument");
  if (kind < 0 || kind >= (int)h->engine.kind_stats.size()) return fail("bad kind");
  *launches = h->engine.kind_stats[kind].launches;
  *ms = h->engine.kind_stats[kind].ms;
  return 0;
}

int spk_set_profile(spk_handle h, int on) {
  if (!h) return fail("null handle");
  h->engine.set_profile(on != 0);
  return 0;
}

int spk_set_profile_mask(spk_handle h, int on, unsigned mask) {
  if (!h) return fail("null handle");
  h->engine.set_profile(on != 0, mask);
  return 0;
}

int spk_device_bytes(spk_handle h, int nrhs, int64_t* bytes) {
  if (!h || !bytes) return fail("null argument");
  return guarded([&]() {
    *bytes = h->engine.device_bytes(nrhs);
    return 0;
  });
}

int spk_release_device(spk_handle h) {
  if (!h) return fail("null handle");
  return guarded([&]() {
    h->engine.release_device();
    return 0;
  });
}

int spk_set_graphs(spk_handle h, int on) {
  if (!h) return fail("null handle");
  h->engine.set_graphs(on != 0);
  return 0;
}

int spk_get_launch_trace(spk_handle h, int32_t* kinds, int64_t* grids, double* ms, int max) {
  if (!h) return fail("null handle");
  const auto& t = h->engine.trace;
  int cnt = (int)std::min<size_t>(t.size(), (size_t)std::max(0, max));
  for (int i = 0; i < cnt; ++i) {
    if (kinds) kinds[i] = t[i].kind;
    if (grids) grids[i] = t[i].grid;
    if (ms) ms[i] = t[i].ms;
  }
  return cnt;
}

int spk_get_launch_levels(spk_handle h, int32_t* levels, int max) {
  if (!h || !levels) return fail("null argument");
  const auto& t = h->engine.trace;
  int cnt = (int)std::min<size_t>(t.size(), (size_t)std::max(0, max));
  for (int i = 0; i < cnt; ++i) levels[i] = t[i].level;
  return cnt;
}

int spk_get_autotune(spk_handle h, int32_t* level, int32_t* split, int32_t* order, double* ms_default,
                     double* ms_best, int max) {
  if (!h || !level || !split || !order || !ms_default || !ms_best) return fail("null argument");
  const auto& t = h->engine.tune_log;
  int cnt = (int)std::min<size_t>(t.size(), (size_t)std::max(0, max));
  for (int i = 0; i < cnt; ++i) {
    level[i] = t[i].level;
    split[i] = t[i].split;
    order[i] = t[i].alt;
    ms_default[i] = t[i].ms0;
    ms_best[i] = t[i].ms;
  }
  return cnt;
}

int spk_get_supernode_levels(spk_handle h, int32_t* level) {
  if (!h || !level) return fail("null argument");
  return guarded([&]() {
    const spk::Plan& p = h->engine.plan();
    if ((int)p.level.size() != p.nsuper) throw std::logic_error("no symbolic factorization");
    std::copy(p.level.begin(), p.level.end(), level);
    return 0;
  });
}

int spk_reset_kernel_stats(spk_handle h) {
  if (!h) return fail("null handle");
  h->engine.reset_stats();
  return 0;
}

int spk_get_supernodal_partition(spk_handle h, int32_t* xsuper) {
  if (!h || !xsuper) return fail("null argument");
  const auto& v = h->engine.sym().xsuper;
  std::memcpy(xsuper, v.data(), v.size() * sizeof(int32_t));
  return 0;
}

int spk_get_sup_membership(spk_handle h, int32_t* supmem) {
  if (!h || !supmem) return fail("null argument");
  const auto& v = h->engine.sym().supMembership;
  std::memcpy(supmem, v.data(), v.size() * sizeof(int32_t));
  return 0;
}

int spk_get_ordering(spk_handle h, int32_t* perm, int32_t* invp) {
  if (!h) return fail("null argument");
  const auto& o = h->engine.sym().ord;
  if (perm) std::memcpy(perm, o.perm.data(), o.perm.size() * sizeof(int32_t));
  if (invp) std::memcpy(invp, o.invp.data(), o.invp.size() * sizeof(int32_t));
  return 0;
}

int spk_get_etree(spk_handle h, int32_t* parent) {
  if (!h || !parent) return fail("null argument");
  const auto& v = h->engine.sym().etree.poparent;
  std::memcpy(parent, v.data(), v.size() * sizeof(int32_t));
  return 0;
}

int spk_get_column_counts(spk_handle h, int32_t* ccu, int32_t* ccr) {
  if (!h) return fail("null argument");
  const auto& s = h->engine.sym();
  if (ccu) std::memcpy(ccu, s.cc_unrelaxed.data(), s.cc_unrelaxed.size() * sizeof(int32_t));
  if (ccr) std::memcpy(ccr, s.cc.data(), s.cc.size() * sizeof(int32_t));
  return 0;
}

int spk_get_structure_size(spk_handle h, int64_t* nnz) {
  if (!h || !nnz) return fail("null argument");
  *nnz = (int64_t)h->engine.sym().lindx.size();
  return 0;
}

int spk_get_structure(spk_handle h, int64_t* xlindx, int32_t* lindx) {
  if (!h) return fail("null argument");
  const auto& s = h->engine.sym();
  if (xlindx) std::memcpy(xlindx, s.xlindx.data(), s.xlindx.size() * sizeof(int64_t));
  if (lindx) std::memcpy(lindx, s.lindx.data(), s.lindx.size() * sizeof(int32_t));
  return 0;
}

int spk_validate_solve_plan(spk_handle h) {
  if (!h) return fail("null handle");
  return guarded([&]() {
    const std::string e = spk::validate_solve_plan(h->engine.plan());
    if (!e.empty()) throw std::logic_error(e);
    return 0;
  });
}

int spk_get_update_merges(spk_handle h, int32_t* deferred_into) {
  if (!h || !deferred_into) return fail("null argument");
  return guarded([&]() {
    const spk::Plan& p = h->engine.plan();
    if ((int)p.deferred_into.size() != p.nsuper) throw std::logic_error("no symbolic factorization");
    std::copy(p.deferred_into.begin(), p.deferred_into.end(), deferred_into);
    return 0;
  });
}

int spk_get_cell_counts(spk_handle h, int64_t* ncells, int64_t* nblocks) {
  if (!h || !ncells || !nblocks) return fail("null argument");
  return guarded([&]() {
    if (!h->cells_built) {
      spk::build_cell_map(h->engine.sym(), h->cells);
      h->cells_built = true;
    }
    *ncells = (int64_t)h->cells.cell_K.size();
    *nblocks = (int64_t)h->cells.block_first_row.size();
    return 0;
  });
}

int spk_get_cells(spk_handle h, int32_t* cK, int32_t* cI, int32_t* cnr, int32_t* cnb, int32_t* bfr,
                  int32_t* bnr) {
  if (!h) return fail("null argument");
  return guarded([&]() {
    if (!h->cells_built) {
      spk::build_cell_map(h->engine.sym(), h->cells);
      h->cells_built = true;
    }
    const auto& c = h->cells;
    size_t nc = c.cell_K.size(), nb = c.block_first_row.size();
    if (cK) std::memcpy(cK, c.cell_K.data(), nc * 4);
    if (cI) std::memcpy(cI, c.cell_I.data(), nc * 4);
    if (cnr) std::memcpy(cnr, c.cell_nrows.data(), nc * 4);
    if (cnb) std::memcpy(cnb, c.cell_nblocks.data(), nc * 4);
    if (bfr) std::memcpy(bfr, c.block_first_row.data(), nb * 4);
    if (bnr) std::memcpy(bnr, c.block_nrows.data(), nb * 4);
    return 0;
  });
}

int spk_get_panel_dims(spk_handle h, int32_t I, int32_t* nrows, int32_t* ncols) {
  if (!h || !nrows || !ncols) return fail("null argument");
  const auto& p = h->engine.plan();
  if (I < 1 || I > p.nsuper) return fail("supernode index out of range");
  *nrows = p.sn[I - 1].r;
  *ncols = p.sn[I - 1].w;
  return 0;
}

int spk_get_factor_panel(spk_handle h, int32_t I, double* out) {
  if (!h || !out) return fail("null argument");
  return guarded([&]() {
    if (!h->engine.factored()) throw std::logic_error("no factor available");
    h->engine.get_panel(I - 1, out);
    return 0;
  });
}

int spk_compute_ordering(int ordering, int32_t n, const int64_t* colptr, const int32_t* rowind,
                         int baseval, int32_t* perm, int32_t* invp) {
  if (!colptr || !rowind || !perm || !invp) return fail("null argument");
  if (n <= 0) return fail("matrix dimension must be positive");
  return guarded([&]() {
    spk::Ordering ord;
    if (ordering == SPK_ORDER_NDBOX) {
      spk::order_ndbox(n, ord);
    } else if (ordering == SPK_ORDER_NDGRID) {
      spk::order_ndgrid(n, ord);
    } else {
      spk::Graph g = spk::expand_symmetric(n, colptr, rowind, baseval);
      spk::set_keep_diag(g, false);
      if (ordering == SPK_ORDER_MMD)
        spk::order_mmd(g, ord);
      else if (ordering == SPK_ORDER_RCM)
        spk::order_rcm(g, ord);
      else if (ordering == SPK_ORDER_AMD)
        spk::order_amd(g, ord);
      else
        throw std::logic_error("spk_compute_ordering: unsupported ordering code");
    }
    std::memcpy(perm, ord.perm.data(), n * sizeof(int32_t));
    std::memcpy(invp, ord.invp.data(), n * sizeof(int32_t));
    return 0;
  });
}

int spk_read_rb(const char* filename, int32_t* n, int64_t* nnz) {
  if (!filename || !n || !nnz) return fail("null argument");
  return guarded([&]() {
    spk::CscMatrix a;
    spk::read_rb(filename, a);
    *n = a.n;
    *nnz = (int64_t)a.rowind.size();
    return 0;
  });
}

int spk_read_rb_data(const char* filename, int64_t* colptr, int32_t* rowind, double* nzval) {
  if (!filename || !colptr || !rowind || !nzval) return fail("null argument");
  return guarded([&]() {
    spk::CscMatrix a;
    spk::read_rb(filename, a);
    std::memcpy(colptr, a.colptr.data(), a.colptr.size() * sizeof(int64_t));
    std::memcpy(rowind, a.rowind.data(), a.rowind.size() * sizeof(int32_t));
    std::memcpy(nzval, a.nzval.data(), a.nzval.size() * sizeof(double));
    return 0;
  });
}

int spk_write_rb(const char* filename, int32_t n, const int64_t* colptr, const int32_t* rowind,
                 const double* nzval, const char* title) {
  if (!filename || !colptr || !rowind || !nzval) return fail("null argument");
  return guarded([&]() {
    spk::CscMatrix a;
    a.n = n;
    a.colptr.assign(colptr, colptr + n + 1);
    int64_t nnz = colptr[n] - colptr[0];
    // stored 1-based as in the Harwell-Boeing format
    int64_t base = colptr[0];
    for (auto& v : a.colptr) v = v - base + 1;
    a.rowind.assign(rowind, rowind + nnz);
    if (base == 0)
      for (auto& r : a.rowind) r += 1;
    a.nzval.assign(nzval, nzval + nnz);
    spk::write_rb(filename, a, title ? title : "sympack_amd", "SPKAMD");
    return 0;
  });
}

int spk_read_matrix(const char* filename, const char* format, int32_t* n, int64_t* nnz) {
  if (!filename || !format || !n || !nnz) return fail("null argument");
  return guarded([&]() {
    spk::CscMatrix a;
    spk::read_matrix(filename, format, a);
    *n = a.n;
    *nnz = (int64_t)a.rowind.size();
    return 0;
  });
}

int spk_read_matrix_data(const char* filename, const char* format, int64_t* colptr, int32_t* rowind,
                         double* nzval) {
  if (!filename || !format || !colptr || !rowind || !nzval) return fail("null argument");
  return guarded([&]() {
    spk::CscMatrix a;
    spk::read_matrix(filename, format, a);
    std::memcpy(colptr, a.colptr.data(), a.colptr.size() * sizeof(int64_t));
    std::memcpy(rowind, a.rowind.data(), a.rowind.size() * sizeof(int32_t));
    std::memcpy(nzval, a.nzval.data(), a.nzval.size() * sizeof(double));
    return 0;
  });
}

int spk_write_matrix(const char* filename, const char* format, int32_t n, const int64_t* colptr,
                     const int32_t* rowind, const double* nzval) {
  if (!filename || !format || !colptr || !rowind || !nzval) return fail("null argument");
  return guarded([&]() {
    spk::CscMatrix a;
    a.n = n;
    const int64_t base = colptr[0];
    const int64_t nnz = colptr[n] - base;
    a.colptr.resize(n + 1);
    for (int32_t i = 0; i <= n; ++i) a.colptr[i] = colptr[i] - base + 1;
    a.rowind.assign(rowind, rowind + nnz);
    if (base == 0)
      for (auto& r : a.rowind) r += 1;
    a.nzval.assign(nzval, nzval + nnz);
    const std::string f = format;
    if (f == "CSC")
      spk::write_csc_binary(filename, a);
    else if (f == "matrix")
      spk::write_matrix_formatted(filename, a);
    else if (f == "HB" || f == "RB" || f == "HARWELL_BOEING")
      spk::write_rb(filename, a, "sympack_amd", "SPKAMD");
    else
      throw std::logic_error("Unknown matrix format.");
    return 0;
  });
}

const char* spk_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
