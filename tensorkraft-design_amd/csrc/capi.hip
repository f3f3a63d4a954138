// extern "C" entry points declared in include/tk_hip.h.
#include "../../include/tk_hip.h"

#include <cstring>
#include <memory>
#include <string>

#include "backend.hpp"
#include "blockops.hpp"
#include "device.hpp"
#include "dmft.hpp"
#include "dmrg.hpp"
#include "heff.hpp"
#include "tdeig.hpp"
#include "tdvp.hpp"
#include "sym.hpp"
#include "truncate.hpp"
#include "tune.hpp"

struct tk_ctx {
  std::unique_ptr<tk::Context> c;
};
struct tk_heff_plan {
  tk_ctx* ctx;
  std::unique_ptr<tk::HeffPlan> p;
};
struct tk_dmrg {
  tk_ctx* ctx;
  std::unique_ptr<tk::DmrgEngine> e;
};
struct tk_broyden {
  tk::BroydenMixer b;
};
struct tk_mlp {
  tk_ctx* ctx;
  std::unique_ptr<tk::Mlp> m;
};

namespace {

thread_local std::string g_error;

void set_error(tk_ctx* ctx, const std::string& msg) {
  g_error = msg;
  if (ctx && ctx->c) ctx->c->last_error = msg;
}

template <typename F>
tk_status guard(tk_ctx* ctx, F&& f) {
  try {
    f();
    return TK_OK;
  } catch (const tk::SymError& e) {
    set_error(ctx, e.what());
    return TK_ERR_SYMMETRY;
  } catch (const tk::ArgError& e) {
    set_error(ctx, e.what());
    return TK_ERR_INVALID_ARG;
  } catch (const tk::CancelledError& e) {
    set_error(ctx, e.what());
    return TK_ERR_CANCELLED;
  } catch (const tk::CapacityError& e) {
    set_error(ctx, e.what());
    return TK_ERR_CAPACITY;
  } catch (const tk::HipError& e) {
    set_error(ctx, e.what());
    return TK_ERR_HIP;
  } catch (const std::bad_alloc&) {
    set_error(ctx, "host allocation failed");
    return TK_ERR_INTERNAL;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return TK_ERR_INTERNAL;
  }
}

void need_ctx(tk_ctx* ctx) {
  if (!ctx || !ctx->c) throw tk::ArgError("null context");
}

tk::Sym to_sym(int32_t s) {
  if (s == TK_SYM_U1) return tk::Sym::U1;
  if (s == TK_SYM_U1U1) return tk::Sym::U1U1;
  throw tk::ArgError("unknown symmetry id");
}
tk::Prec to_prec(int32_t p) {
  if (p == TK_PREC_FP64) return tk::Prec::FP64;
  if (p == TK_PREC_FP32) return tk::Prec::FP32;
  if (p == TK_PREC_BF16X3) return tk::Prec::BF16X3;
  throw tk::ArgError("unknown precision id");
}
tk::Dir to_dir(int32_t d) {
  if (d == TK_DIR_IN) return tk::Dir::In;
  if (d == TK_DIR_OUT) return tk::Dir::Out;
  throw tk::ArgError("unknown leg direction");
}

tk::QIndex to_qindex(tk::Sym sym, const tk_qindex& q) {
  if (q.n_sectors < 0 || (q.n_sectors > 0 && (!q.qn || !q.dims))) throw tk::ArgError("bad tk_qindex");
  std::vector<std::pair<tk::QN, int64_t>> s;
  for (int i = 0; i < q.n_sectors; ++i) {
    tk::QN x{q.qn[2 * i], q.qn[2 * i + 1]};
    (void)tk::pack_qn(sym, x);  // range / component check
    s.push_back({x, q.dims[i]});
  }
  return tk::QIndex(std::move(s));
}

tk::MpoSite to_mpo(tk::Sym sym, const tk_mpo_site* w) {
  if (!w || !w->w) throw tk::ArgError("null MPO site");
  tk::MpoSite m;
  m.phys = to_qindex(sym, w->phys);
  m.wl = to_qindex(sym, w->wl);
  m.wr = to_qindex(sym, w->wr);
  // Callers give QIndex sectors in any order; dense W indices follow the
  // caller's order, so re-map them onto the sorted QIndex order.
  auto perm = [&](const tk_qindex& q, const tk::QIndex& sorted) {
    std::vector<int64_t> p;  // caller full index -> sorted full index
    for (int i = 0; i < q.n_sectors; ++i) {
      int s = sorted.find(tk::QN{q.qn[2 * i], q.qn[2 * i + 1]});
      for (int64_t x = 0; x < q.dims[i]; ++x) p.push_back(sorted.off[s] + x);
    }
    return p;
  };
  auto pp = perm(w->phys, m.phys), pl = perm(w->wl, m.wl), pr = perm(w->wr, m.wr);
  const int64_t d = m.phys.total, a = m.wl.total, b = m.wr.total;
  m.w.assign(size_t(d * d * a * b), 0.0);
  for (int64_t si = 0; si < d; ++si)
    for (int64_t so = 0; so < d; ++so)
      for (int64_t x = 0; x < a; ++x)
        for (int64_t y = 0; y < b; ++y)
          m.w[size_t(((pp[si] * d + pp[so]) * a + pl[x]) * b + pr[y])] = w->w[((si * d + so) * a + x) * b + y];
  return m;
}

tk::BlockTable env_table(tk::Sym sym, const tk_env* e) {
  if (!e) throw tk::ArgError("null environment");
  tk::QIndex bond = to_qindex(sym, e->bond);
  tk::QIndex w = to_qindex(sym, e->w);
  return tk::make_block_table(sym, {bond, w, bond}, {tk::Dir::In, tk::Dir::In, tk::Dir::Out}, tk::QN{});
}

}  // namespace

extern "C" {

const char* tk_version(void) { return "tk-hip 0.1.0 (gfx950)"; }

tk_status tk_ctx_create(int32_t device, tk_ctx** out) {
  return guard(nullptr, [&] {
    if (!out) throw tk::ArgError("null output");
    auto c = std::make_unique<tk_ctx>();
    c->c = std::make_unique<tk::Context>(device);
    *out = c.release();
  });
}

tk_status tk_ctx_destroy(tk_ctx* ctx) {
  return guard(nullptr, [&] { delete ctx; });
}

tk_status tk_ctx_sync(tk_ctx* ctx) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    ctx->c->sync();
  });
}

tk_status tk_ctx_tune(tk_ctx* ctx, int32_t mode) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    if (mode != 0 && mode != 1) throw tk::ArgError("tune mode must be 0 or 1");
    ctx->c->tune_mode = mode;
  });
}

tk_status tk_tune_table(int32_t max, tk_tune_entry* out, int32_t* n_out) {
  return guard(nullptr, [&] {
    if (!n_out || (max > 0 && !out)) throw tk::ArgError("null argument");
    const std::vector<tk::TuneEntry> e = tk::tune_entries();
    *n_out = int32_t(e.size());
    if (int64_t(e.size()) > max) throw tk::CapacityError("tune table: capacity too small");
    for (size_t i = 0; i < e.size(); ++i)
      out[i] = tk_tune_entry{e[i].device, e[i].choice.wg_kind, e[i].choice.tile_min, 0, e[i].sig, e[i].us_best,
                             e[i].us_default};
  });
}

tk_status tk_tune_save(const char* path) {
  return guard(nullptr, [&] {
    if (!path) throw tk::ArgError("null path");
    tk::tune_save(path);
  });
}

tk_status tk_tune_load(const char* path) {
  return guard(nullptr, [&] {
    if (!path) throw tk::ArgError("null path");
    tk::tune_load(path);
  });
}

tk_status tk_tune_clear(void) {
  return guard(nullptr, [&] { tk::tune_clear(); });
}

tk_status tk_ctx_timing(tk_ctx* ctx, int32_t enable) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    ctx->c->timer.enable(enable != 0);
  });
}

tk_status tk_ctx_timing_read(tk_ctx* ctx, int32_t max, tk_kernel_timing* out, int32_t* n_out) {
  return guard(ctx, [&] {
    if (!ctx || !n_out || (max > 0 && !out)) throw tk::ArgError("null argument");
    const std::vector<tk::KernelTimer::Total> t = ctx->c->timer.read();
    if (int64_t(t.size()) > max) throw tk::ArgError("timing: output holds fewer entries than recorded kernels");
    for (size_t i = 0; i < t.size(); ++i) {
      std::memset(out[i].name, 0, sizeof(out[i].name));
      std::strncpy(out[i].name, t[i].name.c_str(), sizeof(out[i].name) - 1);
      out[i].launches = t[i].launches;
      out[i].ms = t[i].ms;
      out[i].bytes = t[i].bytes;
      out[i].flops = t[i].flops;
    }
    *n_out = int32_t(t.size());
  });
}

const char* tk_last_error(tk_ctx* ctx) {
  if (ctx && ctx->c) return ctx->c->last_error.c_str();
  return g_error.c_str();
}

tk_status tk_device_alloc(tk_ctx* ctx, int64_t bytes, void** out) {
  return guard(ctx, [&] {
    if (!ctx || !out || bytes < 0) throw tk::ArgError("bad alloc arguments");
    *out = ctx->c->pool.alloc(size_t(bytes));
  });
}

tk_status tk_device_free(tk_ctx* ctx, void* p) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    ctx->c->sync();
    ctx->c->pool.free(p);
  });
}

tk_status tk_memcpy_h2d(tk_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return guard(ctx, [&] {
    if (!ctx || bytes < 0) throw tk::ArgError("bad copy arguments");
    if (bytes) TK_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyHostToDevice, ctx->c->stream));
    ctx->c->sync();
  });
}

tk_status tk_memcpy_d2h(tk_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return guard(ctx, [&] {
    if (!ctx || bytes < 0) throw tk::ArgError("bad copy arguments");
    if (bytes) TK_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToHost, ctx->c->stream));
    ctx->c->sync();
  });
}

tk_status tk_pack_key(int32_t sym, int32_t rank, const int32_t* qn, uint64_t* key_out) {
  return guard(nullptr, [&] {
    if (rank < 0 || rank > 8 || (rank && !qn) || !key_out) throw tk::ArgError("bad pack arguments");
    tk::QN qs[8];
    for (int i = 0; i < rank; ++i) qs[i] = tk::QN{qn[2 * i], qn[2 * i + 1]};
    *key_out = tk::pack_key(to_sym(sym), qs, rank);
  });
}

tk_status tk_qindex_layout(int32_t sym, const tk_qindex* q, int32_t* qn_out, int64_t* dims_out,
                           int64_t* offsets_out, int64_t* total_out) {
  return guard(nullptr, [&] {
    if (!q) throw tk::ArgError("null qindex");
    tk::QIndex x = to_qindex(to_sym(sym), *q);
    for (int i = 0; i < x.n(); ++i) {
      if (qn_out) {
        qn_out[2 * i] = x.q[i].a;
        qn_out[2 * i + 1] = x.q[i].b;
      }
      if (dims_out) dims_out[i] = x.dim[i];
      if (offsets_out) offsets_out[i] = x.off[i];
    }
    if (total_out) *total_out = x.total;
  });
}

tk_status tk_block_table(int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs, const int32_t* flux,
                         int64_t capacity, int64_t* n_blocks, uint64_t* keys, int64_t* offsets, int32_t* sectors,
                         int64_t* total) {
  return guard(nullptr, [&] {
    if (rank <= 0 || rank > 8 || !legs || !dirs || !flux || !n_blocks) throw tk::ArgError("bad block-table arguments");
    tk::Sym s = to_sym(sym);
    std::vector<tk::QIndex> ql;
    std::vector<tk::Dir> dl;
    for (int i = 0; i < rank; ++i) {
      ql.push_back(to_qindex(s, legs[i]));
      dl.push_back(to_dir(dirs[i]));
    }
    tk::QN fq{flux[0], flux[1]};
    (void)tk::pack_qn(s, fq);
    tk::BlockTable t = tk::make_block_table(s, ql, dl, fq);
    *n_blocks = t.n_blocks();
    if (total) *total = t.total;
    if (t.n_blocks() > capacity) throw tk::CapacityError("output capacity too small");
    for (int64_t b = 0; b < t.n_blocks(); ++b) {
      if (keys) keys[b] = t.keys[b];
      if (offsets) offsets[b] = t.offset[b];
      if (sectors)
        for (int l = 0; l < rank; ++l) sectors[b * rank + l] = t.s(b, l);
    }
  });
}

namespace {
struct Legs {
  tk::Sym sym;
  std::vector<tk::QIndex> q;
  std::vector<tk::Dir> d;
  tk::QN flux;
};
Legs parse_legs(int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs, const int32_t* flux) {
  if (rank <= 0 || rank > 8 || !legs || !dirs || !flux) throw tk::ArgError("bad leg description");
  Legs L;
  L.sym = to_sym(sym);
  for (int i = 0; i < rank; ++i) {
    L.q.push_back(to_qindex(L.sym, legs[i]));
    L.d.push_back(to_dir(dirs[i]));
  }
  L.flux = tk::QN{flux[0], flux[1]};
  (void)tk::pack_qn(L.sym, L.flux);
  return L;
}
}  // namespace

tk_status tk_block_permute(tk_ctx* ctx, int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                           const int32_t* flux, const int32_t* perm, const double* in_dev, int64_t in_elems,
                           double* out_dev, int64_t out_elems) {
  return guard(ctx, [&] {
    if (!ctx || !perm || !in_dev || !out_dev) throw tk::ArgError("null argument");
    Legs L = parse_legs(sym, rank, legs, dirs, flux);
    tk::block_permute(ctx->c.get(), L.sym, L.q, L.d, L.flux, std::vector<int>(perm, perm + rank), in_dev, in_elems,
                      out_dev, out_elems);
  });
}

tk_status tk_block_fuse(tk_ctx* ctx, int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                        const int32_t* flux, int32_t start, int32_t end, const double* in_dev, int64_t in_elems,
                        double* out_dev, int64_t out_elems) {
  return guard(ctx, [&] {
    if (!ctx || !in_dev || !out_dev) throw tk::ArgError("null argument");
    Legs L = parse_legs(sym, rank, legs, dirs, flux);
    tk::block_fuse(ctx->c.get(), L.sym, L.q, L.d, L.flux, start, end, in_dev, in_elems, out_dev, out_elems);
  });
}

tk_status tk_block_split(tk_ctx* ctx, int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                         const int32_t* flux, int32_t start, int32_t end, const double* in_dev, int64_t in_elems,
                         double* out_dev, int64_t out_elems) {
  return guard(ctx, [&] {
    if (!ctx || !in_dev || !out_dev) throw tk::ArgError("null argument");
    Legs L = parse_legs(sym, rank, legs, dirs, flux);
    tk::block_split(ctx->c.get(), L.sym, L.q, L.d, L.flux, start, end, in_dev, in_elems, out_dev, out_elems);
  });
}

tk_status tk_fuse_map(int32_t sym, int32_t n_legs, const tk_qindex* legs, const int32_t* dirs, int64_t capacity,
                      int32_t* n_fused, int32_t* fused_qn, int64_t* fused_dims, int64_t* n_combos,
                      int32_t* combo_fused_sector, int64_t* combo_offset) {
  return guard(nullptr, [&] {
    if (n_legs <= 0 || !legs || !dirs || !n_fused || !n_combos) throw tk::ArgError("bad fuse-map arguments");
    tk::Sym s = to_sym(sym);
    std::vector<tk::QIndex> ql;
    std::vector<tk::Dir> dl;
    for (int i = 0; i < n_legs; ++i) {
      ql.push_back(to_qindex(s, legs[i]));
      dl.push_back(to_dir(dirs[i]));
    }
    tk::FuseMap m = tk::make_fuse_map(ql, dl);
    *n_fused = m.fused.n();
    *n_combos = int64_t(m.combo_offset.size());
    if (m.fused.n() > capacity || int64_t(m.combo_offset.size()) > capacity) throw tk::CapacityError("output capacity too small");
    for (int i = 0; i < m.fused.n(); ++i) {
      if (fused_qn) {
        fused_qn[2 * i] = m.fused.q[i].a;
        fused_qn[2 * i + 1] = m.fused.q[i].b;
      }
      if (fused_dims) fused_dims[i] = m.fused.dim[i];
    }
    for (size_t c = 0; c < m.combo_offset.size(); ++c) {
      if (combo_fused_sector) combo_fused_sector[c] = m.combo_fused_sector[c];
      if (combo_offset) combo_offset[c] = m.combo_offset[c];
    }
  });
}

tk_status tk_gemm(tk_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t rs_a,
                  int64_t cs_a, const double* b, int64_t rs_b, int64_t cs_b, double beta, double* c, int64_t rs_c,
                  int64_t cs_c) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    if ((m > 0 && n > 0 && !c) || (m > 0 && n > 0 && k > 0 && (!a || !b))) throw tk::ArgError("null operand");
    tk::dense_gemm(ctx->c.get(), m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c, cs_c);
  });
}

tk_status tk_gemm_noalloc(tk_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t rs_a,
                          int64_t cs_a, const double* b, int64_t rs_b, int64_t cs_b, double beta, double* c,
                          int64_t rs_c, int64_t cs_c) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    tk::dense_gemm_noalloc(ctx->c.get(), m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c, cs_c);
  });
}

tk_status tk_block_gemm_table(int32_t sym, const tk_block_matrix* a, const tk_block_matrix* b, int64_t capacity,
                              int64_t* n_blocks, uint64_t* keys, int64_t* offsets, int64_t* total) {
  return guard(nullptr, [&] {
    if (!a || !b || !n_blocks) throw tk::ArgError("null argument");
    tk::Sym s = to_sym(sym);
    tk::BlockGemmPlan p(s, to_qindex(s, a->rows), to_qindex(s, a->cols), tk::QN{a->flux[0], a->flux[1]},
                        to_qindex(s, b->rows), to_qindex(s, b->cols), tk::QN{b->flux[0], b->flux[1]});
    *n_blocks = int64_t(p.keys().size());
    if (total) *total = p.total();
    if (int64_t(p.keys().size()) > capacity) throw tk::CapacityError("output capacity too small");
    for (size_t i = 0; i < p.keys().size(); ++i) {
      if (keys) keys[i] = p.keys()[i];
      if (offsets) offsets[i] = p.offsets()[i];
    }
  });
}

tk_status tk_block_gemm(tk_ctx* ctx, int32_t sym, const tk_block_matrix* a, const tk_block_matrix* b, double* c) {
  return guard(ctx, [&] {
    if (!ctx || !a || !b) throw tk::ArgError("null argument");
    tk::Sym s = to_sym(sym);
    tk::BlockGemmPlan p(s, to_qindex(s, a->rows), to_qindex(s, a->cols), tk::QN{a->flux[0], a->flux[1]},
                        to_qindex(s, b->rows), to_qindex(s, b->cols), tk::QN{b->flux[0], b->flux[1]});
    if (p.total() > 0 && (!c || !a->data_dev || !b->data_dev)) throw tk::ArgError("null data");
    p.run(ctx->c.get(), a->data_dev, b->data_dev, c);
  });
}

tk_status tk_spmv(tk_ctx* ctx, int32_t sym, const tk_block_matrix* a, const double* x, double* y) {
  return guard(ctx, [&] {
    if (!ctx || !a || !x || !y) throw tk::ArgError("null argument");
    tk::Sym s = to_sym(sym);
    tk::block_spmv(ctx->c.get(), s, to_qindex(s, a->rows), to_qindex(s, a->cols), tk::QN{a->flux[0], a->flux[1]},
                   a->data_dev, x, y);
  });
}

tk_status tk_svd_truncated(tk_ctx* ctx, int64_t m, int64_t n, const double* a, int64_t max_rank, double cutoff,
                           double* u, double* s, double* vt, int64_t* rank, double* truncation_error) {
  return guard(ctx, [&] {
    if (!ctx || !a || !u || !vt || !rank) throw tk::ArgError("null argument");
    tk::SvdOut o = tk::svd_truncated(ctx->c.get(), m, n, a, max_rank, cutoff, u, vt);
    *rank = o.rank;
    if (s) std::copy(o.s.begin(), o.s.end(), s);
    if (truncation_error) *truncation_error = o.truncation_error;
  });
}

tk_status tk_truncate_svd(tk_ctx* ctx, int64_t rows, int64_t cols, const double* a, int64_t max_bond_dim,
                          double svd_cutoff, int64_t min_bond_dim, double* u, double* s, double* vt,
                          int64_t* bond_dim_new, int64_t* n_discarded, double* truncation_error) {
  return guard(ctx, [&] {
    if (!ctx || !a || !u || !vt || !bond_dim_new) throw tk::ArgError("null argument");
    tk::TruncateOut o =
        tk::truncate_svd(ctx->c.get(), rows, cols, a, max_bond_dim, svd_cutoff, min_bond_dim, u, vt);
    *bond_dim_new = o.bond_dim_new;
    if (s) std::copy(o.s.begin(), o.s.end(), s);
    if (n_discarded) *n_discarded = o.n_discarded;
    if (truncation_error) *truncation_error = o.truncation_error;
  });
}

namespace {
tk::TruncRule make_rule(int64_t max_bond, double cutoff, int64_t min_bond, double degeneracy_tol,
                        double sigma2_floor) {
  if (max_bond < 0 || min_bond < 0) throw tk::ArgError("negative bond dimension");
  if (!(cutoff >= 0) || !(degeneracy_tol >= 0) || !(sigma2_floor >= 0))
    throw tk::ArgError("cutoff, degeneracy tolerance and floor must be non-negative");
  tk::TruncRule r;
  r.max_bond = max_bond;
  r.cutoff = cutoff;
  r.min_bond = min_bond;
  r.degeneracy_tol = degeneracy_tol;
  r.lambda_floor = sigma2_floor;
  return r;
}
}  // namespace

tk_status tk_truncation_select(int32_t n_blocks, const int64_t* counts, const double* sigma, int64_t max_bond,
                               double cutoff, int64_t min_bond, double degeneracy_tol, double sigma2_floor,
                               int64_t* keep, int64_t* n_kept, int64_t* n_discarded, double* truncation_error) {
  return guard(nullptr, [&] {
    if (n_blocks < 0 || (n_blocks > 0 && (!counts || !keep))) throw tk::ArgError("null argument");
    const tk::TruncRule rule = make_rule(max_bond, cutoff, min_bond, degeneracy_tol, sigma2_floor);
    std::vector<std::vector<double>> lam(static_cast<size_t>(n_blocks));
    int64_t at = 0;
    for (int32_t b = 0; b < n_blocks; ++b) {
      if (counts[b] < 0) throw tk::ArgError("negative count");
      if (counts[b] > 0 && !sigma) throw tk::ArgError("null argument");
      for (int64_t j = 0; j < counts[b]; ++j) {
        const double x = sigma[at + j];
        if (!(x >= 0) || !std::isfinite(x)) throw tk::ArgError("singular values must be finite and >= 0");
        if (j > 0 && x > sigma[at + j - 1]) throw tk::ArgError("singular values of a block must descend");
        lam[size_t(b)].push_back(x * x);
      }
      at += counts[b];
    }
    tk::TruncSelection sel = tk::select_truncation(lam, rule);
    for (int32_t b = 0; b < n_blocks; ++b) keep[b] = sel.keep[size_t(b)];
    if (n_kept) *n_kept = sel.kept;
    if (n_discarded) *n_discarded = sel.discarded;
    if (truncation_error) *truncation_error = sel.rel_error();
  });
}

tk_status tk_block_svd_truncate(tk_ctx* ctx, int32_t n_blocks, const int64_t* rows, const int64_t* cols,
                                const double* x, int64_t max_bond, double cutoff, int64_t min_bond,
                                double degeneracy_tol, double sigma2_floor, int64_t* keep, double* sigma,
                                int64_t sigma_capacity, double* u, int64_t u_capacity, double* c,
                                int64_t c_capacity, int64_t* n_kept, int64_t* n_discarded, double* truncation_error) {
  return guard(ctx, [&] {
    if (!ctx || n_blocks < 0 || (n_blocks > 0 && (!rows || !cols || !x || !keep)))
      throw tk::ArgError("null argument");
    tk::Context* cx = ctx->c.get();
    const tk::TruncRule rule = make_rule(max_bond, cutoff, min_bond, degeneracy_tol, sigma2_floor);
    std::vector<tk::GramBlock> gb;
    int64_t xtot = 0;
    for (int32_t b = 0; b < n_blocks; ++b) {
      if (rows[b] < 1 || cols[b] < 1 || rows[b] > tk::kSymEigMaxN || cols[b] > INT32_MAX)
        throw tk::ArgError("block shape outside 1..2048 rows");
      gb.push_back({xtot, int32_t(rows[b]), int32_t(cols[b])});
      xtot += rows[b] * cols[b];
    }
    tk::BlockSvd bs(cx);
    const tk::TruncSelection& sel = bs.run(gb, x, xtot, rule);
    if (!std::isfinite(sel.lambda_max)) throw tk::ArgError("block svd: the blocks are not finite");
    int64_t s_need = 0, u_need = 0, c_need = 0;
    for (int32_t b = 0; b < n_blocks; ++b) {
      s_need += sel.keep[size_t(b)];
      u_need += rows[b] * sel.keep[size_t(b)];
      c_need += sel.keep[size_t(b)] * cols[b];
    }
    if ((s_need > 0 && (!sigma || !u || !c)) || s_need > sigma_capacity || u_need > u_capacity ||
        c_need > c_capacity)
      throw tk::CapacityError("block svd: output capacity too small");
    std::vector<tk::CopyJob> uc;
    tk::ContractList cl;
    cl.flags = tk::kTransA;
    int64_t so = 0, uo = 0, co = 0;
    for (int32_t b = 0; b < n_blocks; ++b) {
      const int64_t k = sel.keep[size_t(b)], m = rows[b], n = cols[b];
      keep[b] = k;
      for (int64_t j = 0; j < k; ++j) sigma[so + j] = std::sqrt(std::max(0.0, bs.lambda()[size_t(b)][size_t(j)]));
      if (k > 0) {
        uc.push_back({bs.u_offset(size_t(b)), uo, int32_t(m), int32_t(k), int32_t(m), int32_t(k), 0, 0});
        // C_b = U_b^T X_b: the kept columns of the n_b x n_b vector matrix, transposed
        cl.begin_out(co, int(k), int(n), int(n));
        cl.begin_term(gb[size_t(b)].x_off, int(n), int(m));
        cl.add_src(bs.u_offset(size_t(b)), int(m), 1.0);
      }
      so += k;
      uo += m * k;
      co += k * n;
    }
    if (!uc.empty() && sel.lambda_max > 0) {
      tk::launch_copies(cx, uc, bs.u(), bs.u_size(), u, u_need);
      cl.finalize();
      cl.validate(bs.u_size(), xtot, c_need);
      tk::DeviceContractList d;
      d.upload(cx, cl);
      tk::launch_contract(cx, d, bs.u(), x, c);
      cx->sync();
    } else if (!uc.empty()) {
      throw tk::ArgError("block svd: all blocks are zero");
    }
    if (n_kept) *n_kept = sel.kept;
    if (n_discarded) *n_discarded = sel.discarded;
    if (truncation_error) *truncation_error = sel.rel_error();
  });
}

tk_status tk_eigh(tk_ctx* ctx, int64_t n, const double* a, double* evals, double* evecs) {
  return guard(ctx, [&] {
    if (!ctx || !a || !evecs) throw tk::ArgError("null argument");
    std::vector<double> w = tk::eigh(ctx->c.get(), n, a, evecs);
    if (evals) std::copy(w.begin(), w.end(), evals);
  });
}

tk_status tk_eigh_lowest(tk_ctx* ctx, int64_t n, const double* a, int64_t k, double* evals, double* evecs) {
  return guard(ctx, [&] {
    if (!ctx || !a || (k > 0 && !evecs)) throw tk::ArgError("null argument");
    std::vector<double> w = tk::eigh_lowest(ctx->c.get(), n, a, k, evecs);
    if (evals) std::copy(w.begin(), w.end(), evals);
  });
}

tk_status tk_qr(tk_ctx* ctx, int64_t m, int64_t n, const double* a, double* q, double* r) {
  return guard(ctx, [&] {
    if (!ctx || !a || !q || !r) throw tk::ArgError("null argument");
    tk::qr(ctx->c.get(), m, n, a, q, r);
  });
}

tk_status tk_fast_quotient(tk_ctx* ctx, const double* a, const double* b, double* q, int64_t n) {
  return guard(ctx, [&] {
    if (!ctx) throw tk::ArgError("null context");
    if (n > 0 && (!a || !b || !q)) throw tk::ArgError("null operand");
    tk::fast_quotient(ctx->c.get(), a, b, q, n);
  });
}

tk_status tk_regularized_svd_inverse(tk_ctx* ctx, int64_t m, int64_t n, int64_t k, const double* s, const double* u,
                                     const double* vt, double delta, double* out) {
  return guard(ctx, [&] {
    if (!ctx || !s || !u || !vt || !out) throw tk::ArgError("null argument");
    tk::regularized_svd_inverse(ctx->c.get(), m, n, k, s, u, vt, delta, out);
  });
}

tk_status tk_bond_schedule_warmup(int64_t d_init, int64_t d_max, int64_t n_warmup_sweeps, int64_t capacity,
                                  int64_t* dims, int64_t* n_dims) {
  return guard(nullptr, [&] {
    if (!dims || !n_dims) throw tk::ArgError("null argument");
    tk::BondSchedule b = tk::BondSchedule::warmup(d_init, d_max, n_warmup_sweeps);
    *n_dims = int64_t(b.dims.size());
    if (int64_t(b.dims.size()) > capacity) throw tk::CapacityError("bond schedule: capacity too small");
    std::copy(b.dims.begin(), b.dims.end(), dims);
  });
}

namespace {
struct CallbackUser {
  tk_matvec_fn fn;
  void* user;
};
void callback_matvec(const double* x, double* y, int, void* u) {
  auto* cb = static_cast<CallbackUser*>(u);
  cb->fn(x, y, cb->user);
}
}  // namespace

tk_status tk_davidson(tk_ctx* ctx, int64_t n, tk_matvec_fn matvec, void* user, const double* diag, double tol,
                      int32_t max_subspace, int32_t max_iter, double* x, double* eigenvalue, double* residual,
                      int32_t* matvecs) {
  return guard(ctx, [&] {
    if (!ctx || !matvec || !x || n <= 0) throw tk::ArgError("bad Davidson arguments");
    tk::DavidsonOpts o;
    o.tol = tol;
    o.max_subspace = max_subspace > 0 ? max_subspace : 60;
    o.max_iter = max_iter > 0 ? max_iter : 1000;
    o.precondition = diag != nullptr;
    CallbackUser cb{matvec, user};
    tk::DavidsonResult r = tk::davidson_generic(ctx->c.get(), n, callback_matvec, &cb, diag, false, x, o);
    if (eigenvalue) *eigenvalue = r.eigenvalue;
    if (residual) *residual = r.residual;
    if (matvecs) *matvecs = r.matvecs;
  });
}

tk_status tk_lanczos(tk_ctx* ctx, int64_t n, tk_matvec_fn matvec, void* user, double tol, int32_t max_krylov,
                     int32_t max_restarts, double* x, double* eigenvalue, double* residual, int32_t* matvecs) {
  return guard(ctx, [&] {
    if (!ctx || !matvec || !x || n <= 0) throw tk::ArgError("bad Lanczos arguments");
    if (max_krylov > 512) throw tk::ArgError("lanczos: max_krylov above 512");
    tk::LanczosOpts o;
    o.tol = tol;
    o.max_krylov = max_krylov > 0 ? max_krylov : 100;
    o.max_restarts = max_restarts >= 0 ? max_restarts : 1000;
    CallbackUser cb{matvec, user};
    tk::DavidsonResult r = tk::lanczos_generic(ctx->c.get(), n, callback_matvec, &cb, false, x, o);
    if (eigenvalue) *eigenvalue = r.eigenvalue;
    if (residual) *residual = r.residual;
    if (matvecs) *matvecs = r.matvecs;
  });
}

tk_status tk_exp_krylov(tk_ctx* ctx, int64_t n, tk_matvec_fn matvec, void* user, const double* v, double alpha,
                        int32_t krylov_dim, double tol, double adaptive_tol, double* out, int32_t* krylov_used) {
  return guard(ctx, [&] {
    if (!ctx || !matvec || !v || !out || n <= 0) throw tk::ArgError("bad exp_krylov arguments");
    CallbackUser cb{matvec, user};
    tk::ExpKrylovResult r =
        tk::exp_krylov_generic(ctx->c.get(), n, callback_matvec, &cb, v, alpha, krylov_dim, tol, out, adaptive_tol);
    if (krylov_used) *krylov_used = r.krylov_dim;
  });
}

tk_status tk_heff_plan_create(tk_ctx* ctx, int32_t prec, int32_t sym, const tk_env* left, const tk_env* right,
                              const tk_mpo_site* wi, const tk_mpo_site* wj, tk_heff_plan** out) {
  return guard(ctx, [&] {
    if (!ctx || !out) throw tk::ArgError("null argument");
    tk::Sym s = to_sym(sym);
    tk::BlockTable el = env_table(s, left), er = env_table(s, right);
    tk::MpoSite mi = to_mpo(s, wi), mj = to_mpo(s, wj);
    auto h = std::make_unique<tk_heff_plan>();
    h->ctx = ctx;
    h->p = std::make_unique<tk::HeffPlan>(ctx->c.get(), to_prec(prec), el, left->data_dev, er, right->data_dev, mi,
                                          mj);
    *out = h.release();
  });
}

tk_status tk_heff_plan_destroy(tk_heff_plan* plan) {
  return guard(nullptr, [&] {
    if (plan) plan->ctx->c->sync();
    delete plan;
  });
}

tk_status tk_heff_psi_table(tk_heff_plan* plan, int64_t capacity, int64_t* n_blocks, uint64_t* keys,
                            int64_t* offsets, int64_t* total) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !n_blocks) throw tk::ArgError("null argument");
    const tk::BlockTable& t = plan->p->psi();
    *n_blocks = t.n_blocks();
    if (total) *total = t.total;
    if (t.n_blocks() > capacity) throw tk::CapacityError("output capacity too small");
    for (int64_t b = 0; b < t.n_blocks(); ++b) {
      if (keys) keys[b] = t.keys[b];
      if (offsets) offsets[b] = t.offset[b];
    }
  });
}

tk_status tk_heff_apply(tk_heff_plan* plan, const double* x, double* y) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !x || !y) throw tk::ArgError("null argument");
    plan->p->apply(x, y);
  });
}

tk_status tk_heff_apply_graph(tk_heff_plan* plan, const double* x, double* y) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !x || !y) throw tk::ArgError("null argument");
    plan->p->apply_graph(x, y);
  });
}

tk_status tk_heff_tiling(tk_heff_plan* plan, int32_t stage, int32_t* wg_kind, int32_t* tile_min) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !wg_kind || !tile_min) throw tk::ArgError("null argument");
    if (stage != 1 && stage != 3) throw tk::ArgError("stage must be 1 or 3");
    const tk::TileChoice c = plan->p->tiling(stage);
    *wg_kind = c.wg_kind;
    *tile_min = c.tile_min;
  });
}

tk_status tk_heff_apply_fp64(tk_heff_plan* plan, const double* x, double* y) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !x || !y) throw tk::ArgError("null argument");
    plan->p->apply_fp64(x, y);
  });
}

tk_status tk_heff_diag(tk_heff_plan* plan, double* d) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !d) throw tk::ArgError("null argument");
    plan->p->diagonal(d);
  });
}

tk_status tk_heff_stats(tk_heff_plan* plan, double* flops, double* bytes, int64_t* counts) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan) throw tk::ArgError("null plan");
    const tk::HeffStats& s = plan->p->stats();
    if (flops) *flops = s.flops;
    if (bytes) *bytes = s.bytes;
    if (counts) {
      counts[0] = s.n_psi_blocks;
      counts[1] = s.dim;
      counts[2] = s.t1_elems;
      counts[3] = s.stage1_items;
      counts[4] = s.stage3_items;
      counts[5] = s.stage3_terms;
    }
  });
}

tk_status tk_heff_stage_stats(tk_heff_plan* plan, double* out) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !out) throw tk::ArgError("null argument");
    const tk::HeffStats& s = plan->p->stats();
    out[0] = s.s1_flops;
    out[1] = s.s1_bytes;
    out[2] = s.s3_flops;
    out[3] = s.s3_bytes;
  });
}

tk_status tk_heff_time(tk_heff_plan* plan, const double* x, double* y, int32_t reps, double* us_apply,
                       double* us_stage1, double* us_stage3) {
  return guard(plan ? plan->ctx : nullptr, [&] {
    if (!plan || !x || !y || reps <= 0) throw tk::ArgError("bad timing arguments");
    tk::Context* c = plan->ctx->c.get();
    hipEvent_t e0, e1, e2;
    TK_HIP(hipEventCreate(&e0));
    TK_HIP(hipEventCreate(&e1));
    TK_HIP(hipEventCreate(&e2));
    double s1 = 0, s3 = 0;
    hipEvent_t a0, a1;
    TK_HIP(hipEventCreate(&a0));
    TK_HIP(hipEventCreate(&a1));
    // whole-apply timing without interleaved events
    TK_HIP(hipEventRecord(a0, c->stream));
    for (int r = 0; r < reps; ++r) plan->p->apply(x, y);
    TK_HIP(hipEventRecord(a1, c->stream));
    TK_HIP(hipEventSynchronize(a1));
    float ms_all = 0;
    TK_HIP(hipEventElapsedTime(&ms_all, a0, a1));
    for (int r = 0; r < reps; ++r) {
      TK_HIP(hipEventRecord(e0, c->stream));
      plan->p->stage1(x);
      TK_HIP(hipEventRecord(e1, c->stream));
      plan->p->stage3(y);
      TK_HIP(hipEventRecord(e2, c->stream));
      TK_HIP(hipEventSynchronize(e2));
      float m1 = 0, m3 = 0;
      TK_HIP(hipEventElapsedTime(&m1, e0, e1));
      TK_HIP(hipEventElapsedTime(&m3, e1, e2));
      s1 += m1;
      s3 += m3;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(a0);
    (void)hipEventDestroy(a1);
    if (us_apply) *us_apply = 1000.0 * ms_all / reps;
    if (us_stage1) *us_stage1 = 1000.0 * s1 / reps;
    if (us_stage3) *us_stage3 = 1000.0 * s3 / reps;
  });
}

tk_status tk_dmrg_create(tk_ctx* ctx, const tk_dmrg_config* cfg, const tk_mpo_site* mpo, tk_dmrg** out) {
  return guard(ctx, [&] {
    if (!ctx || !cfg || !mpo || !out) throw tk::ArgError("null argument");
    if (cfg->n_sites < 2 || cfg->n_schedule < 1 || !cfg->bond_schedule) throw tk::ArgError("bad DMRG config");
    tk::DmrgConfig c;
    c.sym = to_sym(cfg->sym);
    c.prec = to_prec(cfg->prec);
    for (int i = 0; i < cfg->n_sites; ++i) c.mpo.push_back(to_mpo(c.sym, &mpo[i]));
    c.bond_schedule.assign(cfg->bond_schedule, cfg->bond_schedule + cfg->n_schedule);
    for (int64_t d : c.bond_schedule)
      if (d < 1) throw tk::ArgError("bond dimension must be positive");
    c.svd_cutoff = cfg->svd_cutoff;
    c.min_bond_dim = std::max<int64_t>(1, cfg->min_bond_dim);
    c.degeneracy_tol = cfg->degeneracy_tol;
    c.davidson.tol = cfg->davidson_tol;
    c.davidson.max_subspace = cfg->max_subspace > 0 ? cfg->max_subspace : 60;
    c.davidson.max_iter = cfg->max_iter > 0 ? cfg->max_iter : 1000;
    c.davidson.precondition = cfg->precondition != 0;
    c.davidson.switch_rel = cfg->switch_rel > 0 ? cfg->switch_rel : 1e-6;
    c.davidson.low_tail = cfg->low_tail == 0 ? 1000.0 : cfg->low_tail;
    if (cfg->eigensolver != TK_EIG_DAVIDSON && cfg->eigensolver != TK_EIG_LANCZOS)
      throw tk::ArgError("unknown eigensolver");
    if (cfg->max_krylov > 512) throw tk::ArgError("max_krylov above 512");
    c.eigensolver = cfg->eigensolver;
    c.lanczos.tol = c.davidson.tol;
    c.lanczos.max_krylov = cfg->max_krylov > 0 ? cfg->max_krylov : 100;
    c.lanczos.max_restarts = c.davidson.max_iter;
    c.lanczos.switch_rel = c.davidson.switch_rel;
    auto d = std::make_unique<tk_dmrg>();
    d->ctx = ctx;
    d->e = std::make_unique<tk::DmrgEngine>(ctx->c.get(), std::move(c));
    *out = d.release();
  });
}

tk_status tk_dmrg_destroy(tk_dmrg* d) {
  return guard(nullptr, [&] {
    if (d) d->ctx->c->sync();
    delete d;
  });
}

tk_status tk_dmrg_set_state(tk_dmrg* d, const tk_qindex* bonds, const double* const* site_data) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d || !bonds || !site_data) throw tk::ArgError("null argument");
    const int n = d->e->n_sites();
    std::vector<tk::QIndex> b;
    for (int i = 0; i <= n; ++i) b.push_back(to_qindex(d->e->sym(), bonds[i]));
    if (b.front().total != 1 || b.back().total != 1) throw tk::ArgError("boundary bonds must have dimension 1");
    std::vector<std::vector<double>> data(n);
    for (int i = 0; i < n; ++i) {
      tk::BlockTable t = tk::make_block_table(d->e->sym(), {d->e->phys(i), b[i], b[i + 1]},
                                              {tk::Dir::Out, tk::Dir::Out, tk::Dir::In}, tk::QN{});
      if (!site_data[i] && t.total) throw tk::ArgError("null site data");
      data[i].assign(site_data[i], site_data[i] + t.total);
    }
    d->e->set_state(b, data);
  });
}

tk_status tk_dmrg_prepare(tk_dmrg* d) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d) throw tk::ArgError("null argument");
    d->e->prepare();
  });
}

namespace {
void fill_sweep_stats(const tk::SweepStats& s, tk_sweep_stats* out) {
  out->energy = s.energy;
  out->max_truncation_error = s.max_truncation_error;
  out->seconds = s.seconds;
  out->t_plan = s.t_plan;
  out->t_eig = s.t_eig;
  out->t_svd = s.t_svd;
  out->t_env = s.t_env;
  out->max_bond_dim = s.max_bond_dim;
  out->matvecs = s.matvecs;
  out->matvecs_fp64 = s.matvecs_fp64;
  out->steps = s.steps;
  out->unconverged = s.unconverged;
}
}  // namespace

tk_status tk_dmrg_sweep_single_site(tk_dmrg* d, tk_sweep_stats* out) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d) throw tk::ArgError("null argument");
    tk::SweepStats s = d->e->sweep_single_site();
    if (out) fill_sweep_stats(s, out);
  });
}

tk_status tk_dmrg_sweep(tk_dmrg* d, tk_sweep_stats* out) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d) throw tk::ArgError("null argument");
    tk::SweepStats s = d->e->sweep();
    if (out) fill_sweep_stats(s, out);
  });
}

tk_status tk_dmrg_set_cancel_flag(tk_dmrg* d, const int32_t* flag) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d) throw tk::ArgError("null argument");
    d->e->set_cancel_flag(flag);
  });
}

tk_status tk_dmrg_set_subspace_expansion(tk_dmrg* d, double alpha, int64_t max_new, int32_t adaptive) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d) throw tk::ArgError("null argument");
    d->e->set_subspace_expansion(alpha, max_new, adaptive != 0);
  });
}

tk_status tk_dmrg_subspace_expansion_alpha(tk_dmrg* d, double* alpha) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d || !alpha) throw tk::ArgError("null argument");
    *alpha = d->e->subspace_expansion_alpha();
  });
}

tk_status tk_dmrg_apply_site_op(tk_dmrg* d, const double* op, const int32_t* dq, double* norm2) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d || !op || !dq) throw tk::ArgError("null argument");
    const int64_t dim = d->e->phys(0).total;
    std::vector<double> o(op, op + dim * dim);
    const tk::QN q{dq[0], d->e->sym() == tk::Sym::U1 ? 0 : dq[1]};
    const double n2 = d->e->apply_site0_operator(o, q);
    if (norm2) *norm2 = n2;
  });
}

tk_status tk_dmrg_overlap(tk_dmrg* bra, tk_dmrg* ket, double* out) {
  return guard(ket ? ket->ctx : nullptr, [&] {
    if (!bra || !ket || !out) throw tk::ArgError("null argument");
    *out = tk::mps_overlap(*bra->e, *ket->e);
  });
}

tk_status tk_dmrg_tdvp_step(tk_dmrg* d, double tau, int64_t max_bond, int32_t krylov_dim, double krylov_tol,
                            double krylov_adaptive_tol, tk_tdvp_stats* out) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d) throw tk::ArgError("null argument");
    if (krylov_dim <= 0) throw tk::ArgError("tdvp: krylov_dim must be positive");
    tk::TdvpStats s = d->e->tdvp_step(tau, max_bond, krylov_dim, krylov_tol, krylov_adaptive_tol);
    if (out) {
      out->energy = s.energy;
      out->max_truncation_error = s.max_truncation_error;
      out->seconds = s.seconds;
      out->log_norm = s.log_norm;
      out->t_plan = s.t_plan;
      out->t_exp = s.t_exp;
      out->t_svd = s.t_svd;
      out->t_env = s.t_env;
      out->max_bond_dim = s.max_bond_dim;
      out->matvecs = s.matvecs;
      out->pad = 0;
    }
  });
}

tk_status tk_dmrg_steps(tk_dmrg* d, int32_t capacity, int32_t* n, tk_step_stats* out) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d || !n) throw tk::ArgError("null argument");
    const auto& st = d->e->last_steps();
    *n = int32_t(st.size());
    if (int32_t(st.size()) > capacity) throw tk::CapacityError("output capacity too small");
    for (size_t k = 0; k < st.size(); ++k) {
      out[k].site = st[k].site;
      out[k].left_to_right = st[k].left_to_right;
      out[k].iterations = st[k].iterations;
      out[k].matvecs = st[k].matvecs;
      out[k].matvecs_fp64 = st[k].matvecs_fp64;
      out[k].converged = st[k].converged;
      out[k].energy = st[k].energy;
      out[k].truncation_error = st[k].truncation_error;
      out[k].residual = st[k].residual;
      out[k].bond_dim = st[k].bond_dim;
      out[k].dim = st[k].dim;
      out[k].svd_nmax = st[k].svd_nmax;
      out[k].svd_n3 = st[k].svd_n3;
    }
  });
}

tk_status tk_dmrg_bond(tk_dmrg* d, int32_t bond, int32_t capacity, int32_t* n_sectors, int32_t* qn, int64_t* dims) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d || !n_sectors) throw tk::ArgError("null argument");
    if (bond < 0 || bond > d->e->n_sites()) throw tk::ArgError("bond out of range");
    const tk::QIndex& q = d->e->bonds()[bond];
    *n_sectors = q.n();
    if (q.n() > capacity) throw tk::CapacityError("output capacity too small");
    for (int i = 0; i < q.n(); ++i) {
      if (qn) {
        qn[2 * i] = q.q[i].a;
        qn[2 * i + 1] = q.q[i].b;
      }
      if (dims) dims[i] = q.dim[i];
    }
  });
}

tk_status tk_dmrg_site(tk_dmrg* d, int32_t site, int64_t capacity, int64_t* n, double* data) {
  return guard(d ? d->ctx : nullptr, [&] {
    if (!d || !n) throw tk::ArgError("null argument");
    if (site < 0 || site >= d->e->n_sites()) throw tk::ArgError("site out of range");
    std::vector<double> h = d->e->site_data(site);
    *n = int64_t(h.size());
    if (int64_t(h.size()) > capacity) throw tk::CapacityError("output capacity too small");
    if (data) std::copy(h.begin(), h.end(), data);
  });
}


// ---- DMFT impurity pieces ----
tk_status tk_hybridization(tk_ctx* ctx, int32_t geometry, int32_t n_problems, int32_t n_bath,
                           const double* params_dev, int32_t n_z, const double* z_re_dev, const double* z_im_dev,
                           double* out_dev) {
  return guard(ctx, [&] {
    need_ctx(ctx);
    if (geometry != TK_BATH_STAR && geometry != TK_BATH_CHAIN) throw tk::ArgError("unknown bath geometry");
    if (int64_t(n_problems) * n_z > 0 && (!params_dev || !z_re_dev || !z_im_dev || !out_dev))
      throw tk::ArgError("hybridization: null pointer");
    tk::hybridization(ctx->c.get(), tk::BathGeometry(geometry), n_problems, n_bath, params_dev, n_z, z_re_dev,
                      z_im_dev, out_dev);
  });
}

tk_status tk_bath_fit(tk_ctx* ctx, int32_t n_problems, int32_t n_bath, int32_t n_z, const double* wn_dev,
                      const double* delta_dev, double* params_dev, int32_t max_iter, double* cost_host) {
  return guard(ctx, [&] {
    need_ctx(ctx);
    if (n_problems < 0 || max_iter < 0) throw tk::ArgError("bath fit: negative count");
    if (n_problems > 0 && (!wn_dev || !delta_dev || !params_dev || !cost_host))
      throw tk::ArgError("bath fit: null pointer");
    tk::bath_fit(ctx->c.get(), n_problems, n_bath, n_z, wn_dev, delta_dev, params_dev, max_iter, cost_host);
  });
}

tk_status tk_lanczos_tridiagonalize(int32_t n, const double* x, const double* w, int32_t n_bath, double* eps,
                                    double* v) {
  return guard(nullptr, [&] {
    if (n < 0 || n_bath < 0) throw tk::ArgError("discretise: negative size");
    if ((n > 0 && (!x || !w)) || (n_bath > 0 && (!eps || !v))) throw tk::ArgError("discretise: null pointer");
    for (int j = 0; j < n; ++j)
      if (w[j] < 0) throw tk::ArgError("discretise: negative spectral weight");
    tk::lanczos_tridiagonalize(n, x, w, n_bath, eps, v);
  });
}

tk_status tk_broyden_create(int32_t history_depth, tk_broyden** out) {
  return guard(nullptr, [&] {
    if (!out || history_depth < 0) throw tk::ArgError("broyden: bad arguments");
    *out = new tk_broyden{tk::BroydenMixer(history_depth)};
  });
}
tk_status tk_broyden_update(tk_broyden* b, int32_t dim, const double* x, const double* f, double alpha, double* out,
                            int32_t* n_updates) {
  return guard(nullptr, [&] {
    if (!b || dim < 0 || (dim > 0 && (!x || !f || !out))) throw tk::ArgError("broyden: bad arguments");
    const std::vector<double> r = b->b.update(std::vector<double>(x, x + dim), std::vector<double>(f, f + dim), alpha);
    std::copy(r.begin(), r.end(), out);
    if (n_updates) *n_updates = b->b.n_updates();
  });
}
tk_status tk_broyden_destroy(tk_broyden* b) {
  return guard(nullptr, [&] { delete b; });
}

tk_status tk_krylov_tridiag(tk_ctx* ctx, int64_t n, const int64_t* rowptr_dev, const int32_t* col_dev,
                            const double* val_dev, const double* v0_dev, int32_t m, double* alpha, double* beta,
                            int32_t* steps) {
  return guard(ctx, [&] {
    need_ctx(ctx);
    if (!rowptr_dev || !col_dev || !val_dev || !v0_dev || !alpha || !beta || !steps)
      throw tk::ArgError("krylov: null pointer");
    for (int j = 0; j < m; ++j) alpha[j] = beta[j] = 0.0;
    *steps = tk::krylov_tridiag(ctx->c.get(), n, rowptr_dev, col_dev, val_dev, v0_dev, m, alpha, beta);
  });
}

tk_status tk_mlp_create(tk_ctx* ctx, int32_t n_layers, const int32_t* dims, uint64_t seed, tk_mlp** out) {
  return guard(ctx, [&] {
    need_ctx(ctx);
    if (!out || !dims || n_layers < 1) throw tk::ArgError("mlp: bad arguments");
    auto m = std::make_unique<tk_mlp>();
    m->ctx = ctx;
    m->m = std::make_unique<tk::Mlp>(ctx->c.get(), std::vector<int>(dims, dims + n_layers + 1), seed);
    *out = m.release();
  });
}
tk_status tk_mlp_destroy(tk_mlp* m) {
  return guard(nullptr, [&] { delete m; });
}
tk_status tk_mlp_n_params(tk_mlp* m, int64_t* n) {
  return guard(m ? m->ctx : nullptr, [&] {
    if (!m || !n) throw tk::ArgError("mlp: null pointer");
    *n = m->m->n_params();
  });
}
tk_status tk_mlp_get_params(tk_mlp* m, float* host) {
  return guard(m ? m->ctx : nullptr, [&] {
    if (!m || !host) throw tk::ArgError("mlp: null pointer");
    m->m->get_params(host);
  });
}
tk_status tk_mlp_set_params(tk_mlp* m, const float* host) {
  return guard(m ? m->ctx : nullptr, [&] {
    if (!m || !host) throw tk::ArgError("mlp: null pointer");
    m->m->set_params(host);
  });
}
tk_status tk_mlp_forward(tk_mlp* m, int32_t batch, const float* x_dev, float* y_dev) {
  return guard(m ? m->ctx : nullptr, [&] {
    if (!m || !x_dev || !y_dev) throw tk::ArgError("mlp: null pointer");
    m->m->forward(batch, x_dev, y_dev);
  });
}
tk_status tk_mlp_loss_grad(tk_mlp* m, int32_t batch, const float* x_dev, const float* t_dev, float* loss,
                           float* grad_host) {
  return guard(m ? m->ctx : nullptr, [&] {
    if (!m || !x_dev || !t_dev || !loss) throw tk::ArgError("mlp: null pointer");
    *loss = m->m->loss_grad(batch, x_dev, t_dev, grad_host);
  });
}
tk_status tk_mlp_adam(tk_mlp* m, float lr, float beta1, float beta2, float eps) {
  return guard(m ? m->ctx : nullptr, [&] {
    if (!m) throw tk::ArgError("mlp: null pointer");
    m->m->adam_step(lr, beta1, beta2, eps);
  });
}

}  // extern "C"
