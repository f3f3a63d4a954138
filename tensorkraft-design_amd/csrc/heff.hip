// HeffPlan: sector tables -> two grouped-contraction work lists + diagonal.
#include "heff.hpp"

#include "tune.hpp"

#include <cmath>
#include <map>
#include <tuple>

namespace tk {

IndexMap::IndexMap(const QIndex& q) {
  sector.resize(q.total);
  within.resize(q.total);
  for (int s = 0; s < q.n(); ++s)
    for (int64_t x = 0; x < q.dim[s]; ++x) {
      sector[q.off[s] + x] = s;
      within[q.off[s] + x] = int32_t(x);
    }
}

void check_mpo_symmetry(const MpoSite& w) {
  IndexMap P(w.phys), A(w.wl), B(w.wr);
  const int64_t d = w.phys.total;
  if (int64_t(w.w.size()) != d * d * w.wl.total * w.wr.total) throw ArgError("MPO tensor size mismatch");
  for (int64_t si = 0; si < d; ++si)
    for (int64_t so = 0; so < d; ++so)
      for (int64_t a = 0; a < w.wl.total; ++a)
        for (int64_t b = 0; b < w.wr.total; ++b) {
          if (w.at(si, so, a, b) == 0.0) continue;
          QN lhs = fuse(w.phys.q[P.sector[si]], w.wl.q[A.sector[a]]);
          QN rhs = fuse(w.phys.q[P.sector[so]], w.wr.q[B.sector[b]]);
          if (lhs != rhs) throw ArgError("MPO entry violates the charge rule q(s_in)+q(w_l)=q(s_out)+q(w_r)");
        }
}

PairMpo make_pair_mpo(const MpoSite& wi, const MpoSite& wj) {
  if (wi.wr.total != wj.wl.total) throw ArgError("MPO middle bond mismatch");
  PairMpo p;
  p.di = wi.phys.total;
  p.dj = wj.phys.total;
  p.wr_total = wj.wr.total;
  const int64_t wm_total = wi.wr.total;
  // target (soi, soj, wr) -> source (si, sj, wl) -> coefficient, ordered maps
  std::map<std::tuple<int64_t, int64_t, int64_t>, std::map<std::tuple<int64_t, int64_t, int64_t>, double>> acc;
  for (int64_t si = 0; si < p.di; ++si)
    for (int64_t soi = 0; soi < p.di; ++soi)
      for (int64_t wl = 0; wl < wi.wl.total; ++wl)
        for (int64_t wm = 0; wm < wm_total; ++wm) {
          double a = wi.at(si, soi, wl, wm);
          if (a == 0.0) continue;
          for (int64_t sj = 0; sj < p.dj; ++sj)
            for (int64_t soj = 0; soj < p.dj; ++soj)
              for (int64_t wr = 0; wr < p.wr_total; ++wr) {
                double b = wj.at(sj, soj, wm, wr);
                if (b == 0.0) continue;
                acc[{soi, soj, wr}][{si, sj, wl}] += a * b;
              }
        }
  const int64_t nt = p.di * p.dj * p.wr_total;
  p.begin.assign(nt + 1, 0);
  std::vector<std::vector<PairSource>> per(nt);
  for (auto& [t, srcs] : acc) {
    int64_t tid = (std::get<0>(t) * p.dj + std::get<1>(t)) * p.wr_total + std::get<2>(t);
    for (auto& [s, c] : srcs)
      if (c != 0.0) per[tid].push_back({int32_t(std::get<0>(s)), int32_t(std::get<1>(s)), int32_t(std::get<2>(s)), c});
  }
  for (int64_t t = 0; t < nt; ++t) {
    p.begin[t + 1] = p.begin[t] + int64_t(per[t].size());
    for (auto& s : per[t]) p.src.push_back(s);
  }
  return p;
}

HeffPlan::HeffPlan(Context* ctx, Prec prec, const BlockTable& env_l, const double* L, const BlockTable& env_r,
                   const double* R, const MpoSite& wi, const MpoSite& wj)
    : ctx_(ctx), prec_(prec), L_(L), R_(R), l_total_(env_l.total), r_total_(env_r.total) {
  build(env_l, env_r, wi, wj);
}

void HeffPlan::build(const BlockTable& env_l, const BlockTable& env_r, const MpoSite& wi, const MpoSite& wj) {
  if (env_l.rank() != 3 || env_r.rank() != 3) throw ArgError("environments must be rank 3 (bra, w, ket)");
  const QIndex& bondL = env_l.legs[2];
  const QIndex& bondR = env_r.legs[2];
  if (env_l.legs[1].total != wi.wl.total || env_r.legs[1].total != wj.wr.total)
    throw ArgError("environment MPO-bond dimension mismatch");
  check_mpo_symmetry(wi);
  check_mpo_symmetry(wj);
  psi_ = make_block_table(env_l.sym, {wi.phys, wj.phys, bondL, bondR}, {Dir::Out, Dir::Out, Dir::Out, Dir::In}, QN{});
  PairMpo pm = make_pair_mpo(wi, wj);
  IndexMap Pi(wi.phys), Pj(wj.phys), WL(env_l.legs[1]), WR(env_r.legs[1]);
  const QIndex& wlq = env_l.legs[1];
  const QIndex& wrq = env_r.legs[1];

  // T1 registry: (w_l, psi slice) -> T1 index.  psi slice = (block, x_i, x_j).
  std::vector<int64_t> slice_base(psi_.n_blocks() + 1, 0);
  for (int64_t b = 0; b < psi_.n_blocks(); ++b)
    slice_base[b + 1] = slice_base[b] + psi_.dim(b, 0) * psi_.dim(b, 1);
  const int64_t n_slices = slice_base.back();
  std::vector<int32_t> t1_id(size_t(n_slices) * wlq.total, -1);
  struct T1Rec { int64_t off; int32_t m, n, k; int64_t b_off, a_off; int32_t lda; };
  std::vector<T1Rec> t1s;
  int64_t t1_total = 0;

  h3_.clear();
  h3_.flags = kTransB;
  dslices_.clear();
  dterms_.clear();
  for (int64_t bp = 0; bp < psi_.n_blocks(); ++bp) {
    const int si_sec = psi_.s(bp, 0), sj_sec = psi_.s(bp, 1), lp_sec = psi_.s(bp, 2), rp_sec = psi_.s(bp, 3);
    const int64_t dLp = bondL.dim[lp_sec], dRp = bondR.dim[rp_sec];
    const QN qLp = bondL.q[lp_sec], qRp = bondR.q[rp_sec];
    const int64_t dsj = wj.phys.dim[sj_sec];
    for (int64_t xi = 0; xi < wi.phys.dim[si_sec]; ++xi)
      for (int64_t xj = 0; xj < dsj; ++xj) {
        const int64_t poi = wi.phys.off[si_sec] + xi, poj = wj.phys.off[sj_sec] + xj;
        const int64_t y_off = psi_.offset[bp] + (xi * dsj + xj) * dLp * dRp;
        h3_.begin_out(y_off, int(dLp), int(dRp), int(dRp));
        DiagSlice ds{y_off, int32_t(dLp), int32_t(dRp), int32_t(dterms_.size()), 0};
        for (int64_t wr = 0; wr < wrq.total; ++wr) {
          const int swr = WR.sector[wr], iwr = WR.within[wr];
          const QN qR = fuse(qRp, wrq.q[swr]);
          const int r_sec = bondR.find(qR);
          if (r_sec < 0) continue;
          int32_t rsecs[3] = {rp_sec, swr, r_sec};
          const int64_t rb = env_r.find(rsecs);
          if (rb < 0) continue;
          const int64_t dR = bondR.dim[r_sec];
          const int64_t dwr = wrq.dim[swr];
          h3_.begin_term(env_r.offset[rb] + iwr * dR, int(dwr * dR), int(dR));
          const int64_t tid = (poi * pm.dj + poj) * pm.wr_total + wr;
          for (int64_t s = pm.begin[tid]; s < pm.begin[tid + 1]; ++s) {
            const PairSource& src = pm.src[s];
            const int swl = WL.sector[src.wl], iwl = WL.within[src.wl];
            const QN qL = fuse(qLp, wlq.q[swl]);
            const int l_sec = bondL.find(qL);
            if (l_sec < 0) continue;
            int32_t psecs[4] = {Pi.sector[src.si], Pj.sector[src.sj], l_sec, r_sec};
            const int64_t b = psi_.find(psecs);
            if (b < 0) continue;
            int32_t lsecs[3] = {lp_sec, swl, l_sec};
            const int64_t lb = env_l.find(lsecs);
            if (lb < 0) continue;
            const int64_t dL = bondL.dim[l_sec];
            const int64_t xsi = Pi.within[src.si], xsj = Pj.within[src.sj];
            const int64_t slice = slice_base[b] + xsi * psi_.dim(b, 1) + xsj;
            int32_t& id = t1_id[size_t(slice) * wlq.total + src.wl];
            if (id < 0) {
              id = int32_t(t1s.size());
              T1Rec r;
              r.off = t1_total;
              r.m = int32_t(dLp);
              r.n = int32_t(dR);
              r.k = int32_t(dL);
              r.b_off = psi_.offset[b] + (xsi * psi_.dim(b, 1) + xsj) * dL * dR;
              r.a_off = env_l.offset[lb] + iwl * dL;
              r.lda = int32_t(wlq.dim[swl] * dL);
              t1s.push_back(r);
              t1_total += dLp * dR;
            }
            h3_.add_src(t1s[id].off, int(dR), src.coef);
          }
          h3_.drop_empty_term();
        }
        // diagonal: target == source physical indices, w_l and w_r with zero charge
        for (int64_t wr = 0; wr < wrq.total; ++wr) {
          const int swr = WR.sector[wr], iwr = WR.within[wr];
          int32_t rsecs[3] = {rp_sec, swr, rp_sec};
          const int64_t rb = env_r.find(rsecs);
          if (rb < 0) continue;
          const int64_t dwr = wrq.dim[swr];
          const int64_t tid = (poi * pm.dj + poj) * pm.wr_total + wr;
          for (int64_t s = pm.begin[tid]; s < pm.begin[tid + 1]; ++s) {
            const PairSource& src = pm.src[s];
            if (src.si != poi || src.sj != poj) continue;
            const int swl = WL.sector[src.wl], iwl = WL.within[src.wl];
            int32_t lsecs[3] = {lp_sec, swl, lp_sec};
            const int64_t lb = env_l.find(lsecs);
            if (lb < 0) continue;
            const int64_t dwl = wlq.dim[swl];
            DiagTerm t;
            t.l_off = env_l.offset[lb] + iwl * dLp;
            t.l_step = int32_t(dwl * dLp + 1);
            t.r_off = env_r.offset[rb] + iwr * dRp;
            t.r_step = int32_t(dwr * dRp + 1);
            t.coef = src.coef;
            dterms_.push_back(t);
          }
        }
        ds.t_end = int32_t(dterms_.size());
        dslices_.push_back(ds);
      }
  }

  h1_.clear();
  h1_.flags = 0;
  for (const T1Rec& r : t1s) {
    h1_.begin_out(r.off, r.m, r.n, r.n);
    h1_.begin_term(r.b_off, r.n, r.k);
    h1_.add_src(r.a_off, r.lda, 1.0);
  }
  // tiling: the cached measurement for these signatures when tuning is on
  // (a miss is measured at the first apply), else the default
  sig1_ = tile_signature(h1_, int(prec_));
  sig3_ = tile_signature(h3_, int(prec_));
  TileChoice c1, c3;
  tune_pending_ = false;
  if (ctx_->tune_mode)
    tune_pending_ = !(tune_lookup(ctx_->device, sig1_, &c1) && tune_lookup(ctx_->device, sig3_, &c3));
  if (tune_pending_) c1 = c3 = TileChoice{};
  set_tiling(c1, c3, t1_total);
  for (const DiagSlice& s : dslices_)
    for (int t = s.t_begin; t < s.t_end; ++t) {
      const DiagTerm& d = dterms_[t];
      if (d.l_off + int64_t(s.m - 1) * d.l_step >= l_total_ || d.r_off + int64_t(s.n - 1) * d.r_step >= r_total_)
        throw std::out_of_range("heff diag term outside the environments");
    }

  t1_.reset(ctx_, size_t(std::max<int64_t>(t1_total, 1)));

  // diagonal tables in one device allocation
  size_t sbytes = dslices_.size() * sizeof(DiagSlice);
  size_t tbytes = dterms_.size() * sizeof(DiagTerm);
  size_t toff = (sbytes + 255) & ~size_t(255);
  std::vector<char> host(toff + tbytes + 256, 0);
  std::copy_n(reinterpret_cast<const char*>(dslices_.data()), sbytes, host.data());
  std::copy_n(reinterpret_cast<const char*>(dterms_.data()), tbytes, host.data() + toff);
  ddev_.reset(ctx_, host.size());
  ddev_.upload(host.data(), host.size());
  ctx_->sync();  // host staging vectors go out of scope

  stats_.n_psi_blocks = psi_.n_blocks();
  stats_.dim = psi_.total;
  stats_.t1_elems = t1_total;
  stats_.stage1_items = int64_t(h1_.items.size());
  stats_.stage3_items = int64_t(h3_.items.size());
  stats_.stage1_tasks = int64_t(h1_.outs.size());
  stats_.stage3_terms = int64_t(h3_.terms.size());
  stats_.flops = h1_.flops + h3_.flops;
  const double es = prec_ == Prec::FP64 ? 8.0 : 4.0;
  // stage 1 reads L and psi and writes T1; stage 3 reads T1 and R and writes
  // y (FP64 at the boundary).  Each array is counted once per launch.
  stats_.s1_flops = h1_.flops;
  stats_.s3_flops = h3_.flops;
  stats_.s1_bytes = es * (double(l_total_) + double(psi_.total) + double(t1_total));
  stats_.s3_bytes = es * (double(t1_total) + double(r_total_)) + 8.0 * double(psi_.total);
  stats_.s1_bytes_fp64 = 8.0 * (double(l_total_) + double(psi_.total) + double(t1_total));
  stats_.s3_bytes_fp64 = 8.0 * (double(t1_total) + double(r_total_) + double(psi_.total));
  stats_.bytes = stats_.s1_bytes + stats_.s3_bytes;
}

void HeffPlan::set_tiling(const TileChoice& c1, const TileChoice& c3, int64_t t1_total) {
  h1_.wg_kind = c1.wg_kind;
  h1_.tile_min = c1.tile_min;
  h1_.finalize();
  h1_.validate(l_total_, psi_.total, t1_total);
  h3_.wg_kind = c3.wg_kind;
  h3_.tile_min = c3.tile_min;
  h3_.finalize();
  h3_.validate(t1_total, r_total_, psi_.total);
  d1_.upload(ctx_, h1_);
  d3_.upload(ctx_, h3_);
  tiling1_ = c1;
  tiling3_ = c3;
}

void HeffPlan::launch_stage(int stage, const DeviceContractList& d, const double* x, double* y) {
  const bool s1 = stage == 1;
  switch (prec_) {
    case Prec::FP64:
      if (s1) launch_contract(ctx_, d, L_, x, t1_.get());
      else launch_contract(ctx_, d, t1_.get(), R_, y);
      return;
    case Prec::FP32:
      if (s1) launch_contract(ctx_, d, Lf_.get(), xf_.get(), t1f_.get());
      else launch_contract(ctx_, d, t1f_.get(), Rf_.get(), y);
      return;
    case Prec::BF16X3:
      if (s1) launch_contract_bf16x3(ctx_, d, Lf_.get(), xf_.get(), t1f_.get());
      else launch_contract_bf16x3(ctx_, d, t1f_.get(), Rf_.get(), y);
      return;
  }
}

void HeffPlan::tune(const double* x, double* y) {
  tune_pending_ = false;
  if (prec_ != Prec::FP64) {
    ensure_low();
    convert_f64_to_f32(ctx_, x, xf_.get(), psi_.total);
  }
  const std::vector<TileChoice> cands = tile_candidates(int(prec_));
  TileChoice best[2];
  for (int stage : {1, 3}) {
    ContractList& proto = stage == 1 ? h1_ : h3_;
    std::vector<DeviceContractList> dl(cands.size());
    for (size_t i = 0; i < cands.size(); ++i) {
      ContractList l = proto;
      l.wg_kind = cands[i].wg_kind;
      l.tile_min = cands[i].tile_min;
      l.finalize();
      dl[i].upload(ctx_, l);
    }
    ctx_->sync();
    best[stage == 1 ? 0 : 1] = tune_measure(ctx_, stage == 1 ? sig1_ : sig3_, cands,
                                            [&](size_t i) { launch_stage(stage, dl[i], x, y); });
  }
  set_tiling(best[0], best[1], int64_t(t1_.size()));
  drop_graph();
}

void HeffPlan::ensure_low() {
  if (low_ready_) return;
  Lf_.reset(ctx_, size_t(std::max<int64_t>(l_total_, 1)));
  Rf_.reset(ctx_, size_t(std::max<int64_t>(r_total_, 1)));
  convert_f64_to_f32(ctx_, L_, Lf_.get(), l_total_);
  convert_f64_to_f32(ctx_, R_, Rf_.get(), r_total_);
  t1f_.reset(ctx_, t1_.size());
  xf_.reset(ctx_, size_t(std::max<int64_t>(psi_.total, 1)));
  low_ready_ = true;
}

void HeffPlan::patch_graph(const std::vector<int>& args, const void* to) {
  for (size_t i = 0; i < gnodes_.size(); ++i) {
    KernelNode& k = gnodes_[i];
    if (args[i] < 0 || k.slot[args[i]] == reinterpret_cast<uint64_t>(to)) continue;
    k.set(args[i], to);
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(k.func);
    p.gridDim = k.grid;
    p.blockDim = k.block;
    p.sharedMemBytes = k.shmem;
    p.kernelParams = k.args;
    p.extra = nullptr;
    TK_HIP(hipGraphExecKernelNodeSetParams(graph_exec_, ghandles_[i], &p));
  }
}

void HeffPlan::rebind(const double* L, const double* R) {
  L_ = L;
  R_ = R;
  // a low-precision graph reads the FP32 copies, whose buffers stay; an FP64
  // graph's L and R slots are re-pointed
  if (graph_exec_ && prec_ == Prec::FP64) {
    patch_graph(gl_arg_, L_);
    patch_graph(gr_arg_, R_);
  }
  if (low_ready_) {
    convert_f64_to_f32(ctx_, L_, Lf_.get(), l_total_);
    convert_f64_to_f32(ctx_, R_, Rf_.get(), r_total_);
  }
}

void HeffPlan::apply_fp64(const double* x, double* y) {
  if (prec_ == Prec::FP64) {
    apply(x, y);
    return;
  }
  int tk = ctx_->timer.begin(ctx_->stream);
  launch_contract(ctx_, d1_, L_, x, t1_.get());
  ctx_->timer.end(tk, ctx_->stream, "heff_stage1_fp64", stats_.s1_bytes_fp64, stats_.s1_flops, 1);
  tk = ctx_->timer.begin(ctx_->stream);
  launch_contract(ctx_, d3_, t1_.get(), R_, y);
  ctx_->timer.end(tk, ctx_->stream, "heff_stage3_fp64", stats_.s3_bytes_fp64, stats_.s3_flops, 1);
}

void HeffPlan::stage1(const double* x) {
  hipStream_t st = ctx_->stream;
  int tk = -1;
  switch (prec_) {
    case Prec::FP64:
      tk = ctx_->timer.begin(st);
      launch_contract(ctx_, d1_, L_, x, t1_.get());
      ctx_->timer.end(tk, st, "heff_stage1_fp64", stats_.s1_bytes, stats_.s1_flops, 1);
      return;
    case Prec::FP32:
      ensure_low();
      convert_f64_to_f32(ctx_, x, xf_.get(), psi_.total);
      tk = ctx_->timer.begin(st);
      launch_contract(ctx_, d1_, Lf_.get(), xf_.get(), t1f_.get());
      ctx_->timer.end(tk, st, "heff_stage1_fp32", stats_.s1_bytes, stats_.s1_flops, 1);
      return;
    case Prec::BF16X3:
      ensure_low();
      convert_f64_to_f32(ctx_, x, xf_.get(), psi_.total);
      tk = ctx_->timer.begin(st);
      launch_contract_bf16x3(ctx_, d1_, Lf_.get(), xf_.get(), t1f_.get());
      ctx_->timer.end(tk, st, "heff_stage1_bf16x3", stats_.s1_bytes, stats_.s1_flops, 1);
      return;
  }
}

void HeffPlan::stage3(double* y) {
  hipStream_t st = ctx_->stream;
  const int tk = ctx_->timer.begin(st);
  switch (prec_) {
    case Prec::FP64:
      launch_contract(ctx_, d3_, t1_.get(), R_, y);
      ctx_->timer.end(tk, st, "heff_stage3_fp64", stats_.s3_bytes, stats_.s3_flops, 1);
      return;
    case Prec::FP32:
      launch_contract(ctx_, d3_, t1f_.get(), Rf_.get(), y);
      ctx_->timer.end(tk, st, "heff_stage3_fp32", stats_.s3_bytes, stats_.s3_flops, 1);
      return;
    case Prec::BF16X3:
      launch_contract_bf16x3(ctx_, d3_, t1f_.get(), Rf_.get(), y);
      ctx_->timer.end(tk, st, "heff_stage3_bf16x3", stats_.s3_bytes, stats_.s3_flops, 1);
      return;
  }
}

void HeffPlan::apply(const double* x, double* y) {
  if (tune_pending_) tune(x, y);
  stage1(x);
  stage3(y);
}

HeffPlan::~HeffPlan() { drop_graph(); }

void HeffPlan::drop_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  gnodes_.clear();
  ghandles_.clear();
  gx_arg_.clear();
  gy_arg_.clear();
  gl_arg_.clear();
  gr_arg_.clear();
}

void HeffPlan::build_graph(const double* x, double* y) {
  drop_graph();
  std::vector<KernelNode> s1, s3;
  if (prec_ == Prec::FP64) {
    s1 = contract_nodes(ctx_, d1_, L_, x, t1_.get());
    s3 = contract_nodes(ctx_, d3_, t1_.get(), R_, y);
  } else {
    ensure_low();
    gnodes_.push_back(convert_f64_to_f32_node(ctx_, x, xf_.get(), psi_.total));
    gx_arg_.push_back(0);
    gy_arg_.push_back(-1);
    gl_arg_.push_back(-1);
    gr_arg_.push_back(-1);
    if (prec_ == Prec::BF16X3) {
      s1 = contract_bf16x3_nodes(ctx_, d1_, Lf_.get(), xf_.get(), t1f_.get());
      s3 = contract_bf16x3_nodes(ctx_, d3_, t1f_.get(), Rf_.get(), y);
    } else {
      s1 = contract_nodes(ctx_, d1_, Lf_.get(), xf_.get(), t1f_.get());
      s3 = contract_nodes(ctx_, d3_, t1f_.get(), Rf_.get(), y);
    }
  }
  for (const KernelNode& k : s1) {
    gnodes_.push_back(k);
    gx_arg_.push_back(prec_ == Prec::FP64 ? k.b_arg : -1);  // stage 1's B is x
    gy_arg_.push_back(-1);
    gl_arg_.push_back(prec_ == Prec::FP64 ? k.a_arg : -1);  // and its A is L
    gr_arg_.push_back(-1);
  }
  for (const KernelNode& k : s3) {
    gnodes_.push_back(k);
    gx_arg_.push_back(-1);
    gy_arg_.push_back(k.c_arg);  // stage 3's C is y
    gl_arg_.push_back(-1);
    gr_arg_.push_back(prec_ == Prec::FP64 ? k.b_arg : -1);  // and its B is R
  }
  for (KernelNode& k : gnodes_) k.bind();  // slots moved with the vector
  TK_HIP(hipGraphCreate(&graph_, 0));
  hipGraphNode_t prev = nullptr;
  for (KernelNode& k : gnodes_) {
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(k.func);
    p.gridDim = k.grid;
    p.blockDim = k.block;
    p.sharedMemBytes = k.shmem;
    p.kernelParams = k.args;
    p.extra = nullptr;
    hipGraphNode_t h;
    TK_HIP(hipGraphAddKernelNode(&h, graph_, prev ? &prev : nullptr, prev ? 1 : 0, &p));
    ghandles_.push_back(h);
    prev = h;
  }
  TK_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
}

void HeffPlan::apply_graph(const double* x, double* y) {
  if (tune_pending_) tune(x, y);
  if (!graph_exec_) build_graph(x, y);
  for (size_t i = 0; i < gnodes_.size(); ++i) {
    KernelNode& k = gnodes_[i];
    bool changed = false;
    if (gx_arg_[i] >= 0 && k.slot[gx_arg_[i]] != reinterpret_cast<uint64_t>(x)) {
      k.set(gx_arg_[i], x);
      changed = true;
    }
    if (gy_arg_[i] >= 0 && k.slot[gy_arg_[i]] != reinterpret_cast<uint64_t>(y)) {
      k.set(gy_arg_[i], y);
      changed = true;
    }
    if (changed) {
      hipKernelNodeParams p{};
      p.func = const_cast<void*>(k.func);
      p.gridDim = k.grid;
      p.blockDim = k.block;
      p.sharedMemBytes = k.shmem;
      p.kernelParams = k.args;
      p.extra = nullptr;
      TK_HIP(hipGraphExecKernelNodeSetParams(graph_exec_, ghandles_[i], &p));
    }
  }
  const int tk = ctx_->timer.begin(ctx_->stream);
  TK_HIP(hipGraphLaunch(graph_exec_, ctx_->stream));
  ctx_->timer.end(tk, ctx_->stream,
                  prec_ == Prec::FP64 ? "heff_apply_graph_fp64"
                  : prec_ == Prec::FP32 ? "heff_apply_graph_fp32"
                                        : "heff_apply_graph_bf16x3",
                  stats_.s1_bytes + stats_.s3_bytes, stats_.s1_flops + stats_.s3_flops, 1);
}

namespace {
struct DSlice { int64_t y_off; int32_t m, n, t_begin, t_end; };
struct DTerm { int64_t l_off, r_off; int32_t l_step, r_step; double coef; };

__global__ __launch_bounds__(256) void heff_diag_kernel(const DSlice* __restrict__ slices, int n_slices,
                                                        const DTerm* __restrict__ terms,
                                                        const double* __restrict__ L,
                                                        const double* __restrict__ R, double* __restrict__ d) {
  for (int s = blockIdx.x; s < n_slices; s += gridDim.x) {
    const DSlice sl = slices[s];
    const int n_el = sl.m * sl.n;
    for (int e = threadIdx.x; e < n_el; e += blockDim.x) {
      const int a = e / sl.n, b = e - a * sl.n;
      double v = 0.0;
      for (int t = sl.t_begin; t < sl.t_end; ++t) {
        const DTerm tm = terms[t];
        v += tm.coef * L[tm.l_off + int64_t(a) * tm.l_step] * R[tm.r_off + int64_t(b) * tm.r_step];
      }
      d[sl.y_off + e] = v;
    }
  }
}
}  // namespace

void launch_heff_diag(Context* ctx, const void* slices, int n_slices, const void* terms, const double* L,
                      const double* R, double* d) {
  if (n_slices == 0) return;
  int grid = std::min(n_slices, ctx->n_cu * 8);
  hipLaunchKernelGGL(heff_diag_kernel, dim3(grid), dim3(256), 0, ctx->stream,
                     static_cast<const DSlice*>(slices), n_slices, static_cast<const DTerm*>(terms), L, R, d);
  TK_HIP(hipGetLastError());
}

void HeffPlan::diagonal(double* d) {
  static_assert(sizeof(DSlice) == sizeof(DiagSlice), "diag slice layout");
  static_assert(sizeof(DTerm) == sizeof(DiagTerm), "diag term layout");
  size_t sbytes = dslices_.size() * sizeof(DiagSlice);
  size_t toff = (sbytes + 255) & ~size_t(255);
  launch_heff_diag(ctx_, ddev_.get(), int(dslices_.size()), ddev_.get() + toff, L_, R_, d);
}

}  // namespace tk
