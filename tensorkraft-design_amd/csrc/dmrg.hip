// Two-site DMRG engine (see dmrg.hpp).
#include "dmrg.hpp"
#include "truncate.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <numeric>
#include <tuple>

#include "gramsvd.hpp"
#include "tdeig.hpp"
#include "tdvp.hpp"
#include "kernels.hpp"

namespace tk {
namespace {

// Work items of the strided copies: rows [r0, r1) x columns [c0, c1) of one
// job (plain copies; a long row is cut into column chunks) or the 32 x 32
// output tile at (r0, c0) (transposing copies).
struct CopyItem {
  int32_t job, r0, r1, c0, c1, pad;
};
constexpr int kCopyTile = 32;
constexpr int kCopyItemElems = 8192;

// Plain copies as coalesced row segments: a single row with all 256 threads
// along it, several rows four at a time with 64 consecutive columns each.
__global__ __launch_bounds__(256) void copy_rows_kernel(const CopyItem* __restrict__ items, int n_items,
                                                        const CopyJob* __restrict__ jobs,
                                                        const double* __restrict__ src, double* __restrict__ dst) {
  const int ty = threadIdx.x >> 6, tx = threadIdx.x & 63;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const CopyItem ci = items[it];
    const CopyJob c = jobs[ci.job];
    if (ci.r1 - ci.r0 == 1) {
      // four loads in flight per thread before their stores
      const double* s = src + c.src + int64_t(ci.r0) * c.src_ld;
      double* d = dst + c.dst + int64_t(ci.r0) * c.dst_ld;
      int k = ci.c0 + int(threadIdx.x);
      for (; k + 3 * 256 < ci.c1; k += 4 * 256) {
        const double v0 = s[k], v1 = s[k + 256], v2 = s[k + 512], v3 = s[k + 768];
        d[k] = v0;
        d[k + 256] = v1;
        d[k + 512] = v2;
        d[k + 768] = v3;
      }
      for (; k < ci.c1; k += 256) d[k] = s[k];
    } else {
      for (int i = ci.r0 + ty; i < ci.r1; i += 4)
        for (int k = ci.c0 + tx; k < ci.c1; k += 64)
          dst[c.dst + int64_t(i) * c.dst_ld + k] = src[c.src + int64_t(i) * c.src_ld + k];
    }
  }
}

// Transposing copies dst[i][k] = src[k][i]: a 32 x 32 tile is read along i
// (coalesced in src), staged in LDS and written along k (coalesced in dst).
__global__ __launch_bounds__(256) void copy_tiles_kernel(const CopyItem* __restrict__ items, int n_items,
                                                         const CopyJob* __restrict__ jobs,
                                                         const double* __restrict__ src, double* __restrict__ dst) {
  __shared__ double t[kCopyTile][kCopyTile + 1];
  const int tx = threadIdx.x & (kCopyTile - 1), ty = threadIdx.x >> 5;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const CopyItem ci = items[it];
    const CopyJob c = jobs[ci.job];
    for (int r = ty; r < kCopyTile; r += 8) {
      const int k = ci.c0 + r, i = ci.r0 + tx;
      if (k < c.n && i < c.m) t[r][tx] = src[c.src + int64_t(k) * c.src_ld + i];
    }
    __syncthreads();
    for (int r = ty; r < kCopyTile; r += 8) {
      const int i = ci.r0 + r, k = ci.c0 + tx;
      if (i < c.m && k < c.n) dst[c.dst + int64_t(i) * c.dst_ld + k] = t[tx][r];
    }
    __syncthreads();
  }
}

__global__ void fill_kernel2(double* __restrict__ p, double v, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    p[j] = v;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool same_qindex(const QIndex& a, const QIndex& b) {
  if (a.n() != b.n()) return false;
  for (int i = 0; i < a.n(); ++i)
    if (a.q[i] != b.q[i] || a.dim[i] != b.dim[i]) return false;
  return true;
}

// Upload a work list and launch one contraction.
template <typename TA, typename TB, typename TC>
void run_list(Context* ctx, ContractList& cl, int64_t a_n, int64_t b_n, int64_t c_n, const TA* A, const TB* B,
              TC* C) {
  cl.finalize();
  cl.validate(a_n, b_n, c_n);
  DeviceContractList d;
  d.upload(ctx, cl);
  launch_contract(ctx, d, A, B, C);
}

}  // namespace

void launch_copies(Context* ctx, const std::vector<CopyJob>& jobs, const double* src, int64_t src_n, double* dst,
                   int64_t dst_n) {
  if (jobs.empty()) return;
  for (const CopyJob& c : jobs) {
    if (c.m <= 0 || c.n <= 0) throw std::out_of_range("copy: empty job");
    const int64_t s_end = c.trans ? c.src + int64_t(c.n - 1) * c.src_ld + c.m : c.src + int64_t(c.m - 1) * c.src_ld + c.n;
    const int64_t d_end = c.dst + int64_t(c.m - 1) * c.dst_ld + c.n;
    if (c.src < 0 || c.dst < 0 || s_end > src_n || d_end > dst_n || c.dst_ld < c.n)
      throw std::out_of_range("copy: job outside its buffers");
  }
  std::vector<CopyItem> rows, tiles;
  double row_elems = 0, tile_elems = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const CopyJob& c = jobs[j];
    (c.trans ? tile_elems : row_elems) += double(c.m) * double(c.n);
    if (c.trans) {
      for (int i0 = 0; i0 < c.m; i0 += kCopyTile)
        for (int k0 = 0; k0 < c.n; k0 += kCopyTile) tiles.push_back({int32_t(j), i0, i0 + 1, k0, k0 + 1, 0});
    } else if (c.n >= kCopyItemElems / 2) {
      for (int r = 0; r < c.m; ++r)
        for (int k0 = 0; k0 < c.n; k0 += kCopyItemElems / 2)
          rows.push_back({int32_t(j), r, r + 1, k0, std::min(c.n, k0 + kCopyItemElems / 2), 0});
    } else {
      const int nr = std::max(1, kCopyItemElems / c.n);
      for (int r0 = 0; r0 < c.m; r0 += nr) rows.push_back({int32_t(j), r0, std::min(c.m, r0 + nr), 0, c.n, 0});
    }
  }
  DBuf<CopyJob> d(ctx, jobs.size());
  d.upload(jobs.data(), jobs.size());
  DBuf<CopyItem> di(ctx, rows.size() + tiles.size());
  if (!rows.empty()) di.upload(rows.data(), rows.size());
  if (!tiles.empty())
    ctx->upload(di.get() + rows.size(), tiles.data(), tiles.size() * sizeof(CopyItem));
  const size_t cap = size_t(ctx->n_cu) * 8;
  // algorithmic bytes: each element read once and written once
  if (!rows.empty()) {
    const int tk = ctx->timer.begin(ctx->stream);
    hipLaunchKernelGGL(copy_rows_kernel, dim3(unsigned(std::min(rows.size(), cap))), dim3(256), 0, ctx->stream,
                       di.get(), int(rows.size()), d.get(), src, dst);
    ctx->timer.end(tk, ctx->stream, "copy_rows", 16.0 * row_elems, 0.0, 1);
  }
  if (!tiles.empty()) {
    const int tk = ctx->timer.begin(ctx->stream);
    hipLaunchKernelGGL(copy_tiles_kernel, dim3(unsigned(std::min(tiles.size(), cap))), dim3(256), 0, ctx->stream,
                       di.get() + rows.size(), int(tiles.size()), d.get(), src, dst);
    ctx->timer.end(tk, ctx->stream, "copy_tiles", 16.0 * tile_elems, 0.0, 1);
  }
  TK_HIP(hipGetLastError());
  ctx->sync();
}

void launch_fill(Context* ctx, double* p, double v, int64_t n) {
  if (n <= 0) return;
  int grid = int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, int64_t(ctx->n_cu) * 16)));
  hipLaunchKernelGGL(fill_kernel2, dim3(grid), dim3(256), 0, ctx->stream, p, v, n);
  TK_HIP(hipGetLastError());
}

DmrgEngine::DmrgEngine(Context* ctx, DmrgConfig cfg) : ctx_(ctx), cfg_(std::move(cfg)) {
  n_ = int(cfg_.mpo.size());
  if (n_ < 2) throw ArgError("DMRG needs at least two sites");
  if (cfg_.mpo.front().wl.total != 1 || cfg_.mpo.back().wr.total != 1)
    throw ArgError("boundary MPO bonds must have dimension 1");
  for (int i = 0; i < n_; ++i) {
    check_mpo_symmetry(cfg_.mpo[i]);
    if (i + 1 < n_ && !same_qindex(cfg_.mpo[i].wr, cfg_.mpo[i + 1].wl))
      throw ArgError("MPO bond " + std::to_string(i + 1) + " differs between neighbouring sites");
  }
  if (cfg_.bond_schedule.empty()) throw ArgError("empty bond-dimension schedule");
}

QIndex DmrgEngine::trivial_bond(QN q) const { return QIndex({{q, 1}}); }

void DmrgEngine::set_state(const std::vector<QIndex>& bonds, const std::vector<std::vector<double>>& data) {
  if (int(bonds.size()) != n_ + 1 || int(data.size()) != n_) throw ArgError("state size mismatch");
  bonds_ = bonds;
  site_.clear();
  site_.resize(n_);
  for (int i = 0; i < n_; ++i) {
    site_[i].t = make_block_table(cfg_.sym, {cfg_.mpo[i].phys, bonds_[i], bonds_[i + 1]}, {Dir::Out, Dir::Out, Dir::In},
                                  QN{});
    if (int64_t(data[i].size()) != site_[i].t.total)
      throw ArgError("site " + std::to_string(i) + " data size " + std::to_string(data[i].size()) + " != table " +
                     std::to_string(site_[i].t.total));
    site_[i].d.reset(ctx_, size_t(std::max<int64_t>(1, site_[i].t.total)));
    site_[i].d.upload(data[i].data(), data[i].size());
  }
  ctx_->sync();
  L_.clear();
  R_.clear();
  L_.resize(n_ + 1);
  R_.resize(n_ + 1);
  sweeps_ = 0;
  centre_ = -1;
}

std::vector<double> DmrgEngine::site_data(int i) const {
  std::vector<double> h(size_t(site_[i].t.total));
  site_[i].d.download(h.data(), h.size());
  ctx_->sync();
  return h;
}

DmrgEngine::Env DmrgEngine::boundary_env(const QIndex& bond, const QIndex& w) {
  Env e;
  e.t = make_block_table(cfg_.sym, {bond, w, bond}, {Dir::In, Dir::In, Dir::Out}, QN{});
  if (e.t.total != 1) throw ArgError("boundary environment must be 1x1x1");
  e.d.reset(ctx_, 1);
  launch_fill(ctx_, e.d.get(), 1.0, 1);
  return e;
}

void DmrgEngine::prepare() {
  if (site_.empty()) throw ArgError("set_state first");
  // right-canonicalise with exact (rank-only) splits, centre ends at site 0
  for (int i = n_ - 2; i >= 0; --i) {
    BlockTable psi = make_block_table(cfg_.sym, {cfg_.mpo[i].phys, cfg_.mpo[i + 1].phys, bonds_[i], bonds_[i + 2]},
                                      {Dir::Out, Dir::Out, Dir::Out, Dir::In}, QN{});
    DBuf<double> x(ctx_, size_t(std::max<int64_t>(1, psi.total)));
    contract_two_site(i, psi, x.get());
    StepStats st;
    split_two_site(i, false, psi, x.get(), INT64_MAX, 0.0, kLambdaZero, &st);
  }
  L_[0] = boundary_env(bonds_[0], cfg_.mpo[0].wl);
  R_[n_] = boundary_env(bonds_[n_], cfg_.mpo[n_ - 1].wr);
  for (int i = n_ - 1; i >= 2; --i) grow_right(i);
  ctx_->sync();
  centre_ = 0;
}

void DmrgEngine::centre_to_site0() {
  if (centre_ < 0) throw ArgError("prepare() first");
  const int c = centre_;
  for (int j = c - 1; j >= 0; --j) {
    const BlockTable pair = make_block_table(cfg_.sym, {cfg_.mpo[j].phys, cfg_.mpo[j + 1].phys, bonds_[j], bonds_[j + 2]},
                                             {Dir::Out, Dir::Out, Dir::Out, Dir::In}, QN{});
    DBuf<double> th(ctx_, size_t(std::max<int64_t>(1, pair.total)));
    contract_two_site(j, pair, th.get());
    StepStats sp;
    split_two_site(j, false, pair, th.get(), INT64_MAX, 0.0, kLambdaZero, &sp);
    grow_right(j + 1);
  }
  if (c == 0) grow_right(1);
  ctx_->sync();
}

void DmrgEngine::check_cancel() {
  if (cancel_ && __atomic_load_n(cancel_, __ATOMIC_RELAXED) != 0) {
    cur_ = nullptr;
    centre_ = -1;  // prepare() again before sweeping
    throw CancelledError("computation cancelled");
  }
}

SweepStats DmrgEngine::sweep() {
  if (centre_ < 0) throw ArgError("sweep: prepare the state first");
  SweepStats s;
  cur_ = &s;
  steps_.clear();
  const double t0 = now();
  for (int i = 0; i + 1 < n_; ++i) {
    check_cancel();
    steps_.push_back(step(i, true));
  }
  for (int i = n_ - 1; i >= 1; --i) {
    check_cancel();
    steps_.push_back(step(i, false));
  }
  ctx_->sync();
  s.seconds = now() - t0;
  for (const StepStats& st : steps_) {
    if (st.iterations == 0) continue;
    ++s.steps;
    s.max_truncation_error = std::max(s.max_truncation_error, st.truncation_error);
    s.max_bond_dim = std::max(s.max_bond_dim, st.bond_dim);
    s.matvecs += st.matvecs;
    s.matvecs_fp64 += st.matvecs_fp64;
    if (!st.converged) ++s.unconverged;
  }
  // sweep energy = the last step's eigenvalue (sweep/mod.rs:301, 313-314)
  for (auto it = steps_.rbegin(); it != steps_.rend(); ++it)
    if (it->iterations > 0) {
      energy_ = it->energy;
      break;
    }
  s.energy = energy_;
  ++sweeps_;
  cur_ = nullptr;
  return s;
}

namespace {
bool same_layout(const BlockTable& a, const BlockTable& b) {
  if (a.total != b.total || a.keys != b.keys || a.offset != b.offset || a.sec != b.sec || a.legs.size() != b.legs.size())
    return false;
  for (size_t k = 0; k < a.legs.size(); ++k)
    if (a.legs[k].q != b.legs[k].q || a.legs[k].dim != b.legs[k].dim) return false;
  return true;
}
}  // namespace

HeffPlan& DmrgEngine::two_site_plan(int i) {
  if (plan_cache_.size() < size_t(n_)) plan_cache_.resize(size_t(n_));
  CachedPlan& c = plan_cache_[size_t(i)];
  const Env& l = L_[i];
  const Env& r = R_[i + 2];
  if (c.plan && same_layout(c.l, l.t) && same_layout(c.r, r.t)) {
    c.plan->rebind(l.d.get(), r.d.get());
    return *c.plan;
  }
  c.plan.reset();  // release its device buffers before the new plan allocates
  c.plan = std::make_unique<HeffPlan>(ctx_, cfg_.prec, l.t, l.d.get(), r.t, r.d.get(), cfg_.mpo[i], cfg_.mpo[i + 1]);
  c.l = l.t;
  c.r = r.t;
  return *c.plan;
}

StepStats DmrgEngine::step(int i, bool lr) {
  StepStats st;
  st.site = i;
  st.left_to_right = lr;
  if (i + 1 >= n_) {  // right-to-left sweep's first site: no bond to its right (sweep/mod.rs:354-364)
    st.energy = energy_;
    st.bond_dim = bonds_[std::min(i, n_)].total;
    return st;
  }
  const int64_t max_bond =
      cfg_.bond_schedule[std::min<size_t>(size_t(sweeps_), cfg_.bond_schedule.size() - 1)];
  double t0 = now();
  HeffPlan& plan = two_site_plan(i);
  const BlockTable& psi = plan.psi();
  DBuf<double> x(ctx_, size_t(std::max<int64_t>(1, psi.total)));
  contract_two_site(i, psi, x.get());
  ctx_->sync();
  double t1 = now();
  DavidsonResult dr = cfg_.eigensolver == 1 ? lanczos(ctx_, plan, x.get(), cfg_.lanczos)
                                            : davidson(ctx_, plan, x.get(), cfg_.davidson);
  double t2 = now();
  st.energy = dr.eigenvalue;
  st.iterations = dr.iterations;
  st.matvecs = dr.matvecs;
  st.matvecs_fp64 = dr.matvecs_fp64;
  st.converged = dr.converged;
  st.residual = dr.residual;
  const double floor = std::max(cfg_.svd_cutoff * cfg_.svd_cutoff, kLambdaZero);
  split_two_site(i, lr, psi, x.get(), max_bond, cfg_.svd_cutoff, floor, &st);
  double t3 = now();
  if (lr) {
    grow_left(i);
  } else {
    grow_right(i + 1);
  }
  ctx_->sync();
  double t4 = now();
  if (cur_) {
    cur_->t_plan += t1 - t0;
    cur_->t_eig += t2 - t1;
    cur_->t_svd += t3 - t2;
    cur_->t_env += t4 - t3;
  }
  return st;
}

struct DmrgEngine::OneSite {
  MpoSite id;
  std::unique_ptr<HeffPlan> plan;
  std::vector<CopyJob> to_psi, to_site;
};

DmrgEngine::OneSite DmrgEngine::one_site_plan(int i, Prec prec) {
  // a trivial second site: one state of charge 0 and W = identity on the MPO
  // bond, so that psi(s, 0, a_L, a_R) is the site tensor (s, a_L, a_R) block
  // for block
  const MpoSite& W = cfg_.mpo[i];
  OneSite os;
  os.id.phys = QIndex({{QN{}, 1}});
  os.id.wl = W.wr;
  os.id.wr = W.wr;
  os.id.w.assign(size_t(W.wr.total) * size_t(W.wr.total), 0.0);
  for (int64_t a = 0; a < W.wr.total; ++a) os.id.w[size_t(a) * size_t(W.wr.total) + size_t(a)] = 1.0;
  os.plan = std::make_unique<HeffPlan>(ctx_, prec, L_[i].t, L_[i].d.get(), R_[i + 1].t, R_[i + 1].d.get(), W, os.id);
  const BlockTable& psi = os.plan->psi();
  const Site& A = site_[i];
  for (int64_t b = 0; b < psi.n_blocks(); ++b) {
    int32_t secs[3] = {psi.s(b, 0), psi.s(b, 2), psi.s(b, 3)};
    const int64_t ab = A.t.find(secs);
    if (ab < 0) throw std::logic_error("one-site: site block missing from the one-site vector");
    const int32_t rows = int32_t(W.phys.dim[secs[0]] * bonds_[i].dim[secs[1]]);
    const int32_t cols = int32_t(bonds_[i + 1].dim[secs[2]]);
    os.to_psi.push_back({A.t.offset[ab], psi.offset[b], rows, cols, cols, cols, 0, 0});
    os.to_site.push_back({psi.offset[b], A.t.offset[ab], rows, cols, cols, cols, 0, 0});
  }
  if (int64_t(os.to_psi.size()) != A.t.n_blocks()) throw std::logic_error("one-site: block sets differ");
  return os;
}

double DmrgEngine::evolve_one_site(int i, double alpha, int krylov_dim, double tol, double adaptive_tol,
                                   TdvpStats* ts) {
  const double t0 = now();
  OneSite os = one_site_plan(i, Prec::FP64);
  const double t1 = now();
  const BlockTable& psi = os.plan->psi();
  Site& A = site_[i];
  DBuf<double> x(ctx_, size_t(std::max<int64_t>(1, psi.total))), y(ctx_, size_t(std::max<int64_t>(1, psi.total)));
  launch_copies(ctx_, os.to_psi, A.d.get(), A.t.total, x.get(), psi.total);
  const ExpKrylovResult r = exp_krylov(ctx_, *os.plan, x.get(), alpha, krylov_dim, tol, y.get(), adaptive_tol);
  ts->matvecs += r.matvecs;
  if (!(r.norm_out > 0) || !(r.norm_in > 0)) throw ArgError("tdvp: the one-site tensor vanished");
  scale(ctx_, y.get(), 1.0 / r.norm_out, psi.total);
  launch_copies(ctx_, os.to_site, y.get(), psi.total, A.d.get(), A.t.total);
  ts->t_plan += t1 - t0;
  ts->t_exp += now() - t1;
  return std::log(r.norm_out) - std::log(r.norm_in);
}

namespace {
// x . y of two device vectors, on the host
double dot_host(Context* ctx, const double* x, const double* y, int64_t n) {
  DBuf<double> o(ctx, 1);
  dot_many(ctx, x, n, 1, y, n, o.get());
  double h = 0;
  o.download(&h, 1);
  ctx->sync();
  return h;
}
}  // namespace

StepStats DmrgEngine::step_single_site(int i, bool lr) {
  StepStats st;
  st.site = i;
  st.left_to_right = lr;
  double t0 = now();
  OneSite os = one_site_plan(i, cfg_.prec);
  const BlockTable& psi = os.plan->psi();
  Site& A = site_[i];
  DBuf<double> x(ctx_, size_t(std::max<int64_t>(1, psi.total)));
  launch_copies(ctx_, os.to_psi, A.d.get(), A.t.total, x.get(), psi.total);
  double t1 = now();
  DavidsonResult dr = cfg_.eigensolver == 1 ? lanczos(ctx_, *os.plan, x.get(), cfg_.lanczos)
                                            : davidson(ctx_, *os.plan, x.get(), cfg_.davidson);
  launch_copies(ctx_, os.to_site, x.get(), psi.total, A.d.get(), A.t.total);
  double t2 = now();
  st.energy = dr.eigenvalue;
  st.iterations = dr.iterations;
  st.matvecs = dr.matvecs;
  st.matvecs_fp64 = dr.matvecs_fp64;
  st.converged = dr.converged;
  st.residual = dr.residual;
  st.dim = psi.total;
  // move the centre: untruncated split of the site with its neighbour, or
  // with subspace expansion a truncated split of the perturbed pair (only
  // where a one-site solve follows: not into site 0 at the end of the sweep)
  const int j = lr ? i : i - 1;  // left site of the pair
  st.bond_dim = bonds_[lr ? i + 1 : i].total;
  if (lr ? i + 1 < n_ : i > 0) {
    const bool expand = expand_alpha_ > 0 && (lr || i > 1);
    const BlockTable pair = make_block_table(cfg_.sym, {cfg_.mpo[j].phys, cfg_.mpo[j + 1].phys, bonds_[j], bonds_[j + 2]},
                                             {Dir::Out, Dir::Out, Dir::Out, Dir::In}, QN{});
    DBuf<double> th(ctx_, size_t(std::max<int64_t>(1, pair.total)));
    contract_two_site(j, pair, th.get());
    StepStats sp;
    if (expand) {
      HeffPlan plan(ctx_, Prec::FP64, L_[j].t, L_[j].d.get(), R_[j + 2].t, R_[j + 2].d.get(), cfg_.mpo[j],
                    cfg_.mpo[j + 1]);
      if (plan.psi().total != pair.total || plan.psi().n_blocks() != pair.n_blocks())
        throw std::logic_error("expansion: the plan's two-site table differs from the pair's");
      DBuf<double> r(ctx_, size_t(std::max<int64_t>(1, pair.total)));
      plan.apply_fp64(th.get(), r.get());
      // r_perp = r - (theta.r / theta.theta) theta, theta += alpha |theta| r_perp / |r_perp|
      const double tt = dot_host(ctx_, th.get(), th.get(), pair.total);
      const double tr = dot_host(ctx_, th.get(), r.get(), pair.total);
      if (tt > 0) {
        axpy(ctx_, -tr / tt, th.get(), r.get(), pair.total);
        const double rr = dot_host(ctx_, r.get(), r.get(), pair.total);
        if (rr > 1e-30 * tt * std::max(1.0, tr * tr / (tt * tt)))
          axpy(ctx_, expand_alpha_ * std::sqrt(tt / rr), r.get(), th.get(), pair.total);
      }
      const int64_t cur = bonds_[j + 1].total;
      const int64_t sched =
          cfg_.bond_schedule[std::min<size_t>(size_t(sweeps_), cfg_.bond_schedule.size() - 1)];
      const int64_t cap = expand_max_new_ > 0 ? std::min(sched, cur + expand_max_new_) : sched;
      const double floor = std::max(cfg_.svd_cutoff * cfg_.svd_cutoff, kLambdaZero);
      split_two_site(j, lr, pair, th.get(), std::max(cap, cur), cfg_.svd_cutoff, floor, &sp);
    } else {
      split_two_site(j, lr, pair, th.get(), INT64_MAX, 0.0, kLambdaZero, &sp);
    }
    st.truncation_error = sp.truncation_error;
    st.bond_dim = sp.bond_dim;
    st.svd_nmax = sp.svd_nmax;
    st.svd_n3 = sp.svd_n3;
  }
  double t3 = now();
  if (lr && i + 1 < n_) grow_left(i);
  if (!lr && i > 0) grow_right(i);
  ctx_->sync();
  double t4 = now();
  if (cur_) {
    cur_->t_plan += t1 - t0;
    cur_->t_eig += t2 - t1;
    cur_->t_svd += t3 - t2;
    cur_->t_env += t4 - t3;
  }
  return st;
}

void DmrgEngine::set_subspace_expansion(double alpha, int64_t max_new, bool adaptive) {
  if (!(alpha >= 0) || !std::isfinite(alpha)) throw ArgError("subspace expansion: alpha must be finite and >= 0");
  expand_alpha_ = alpha;
  expand_max_new_ = max_new;
  expand_adaptive_ = adaptive;
  expand_have_prev_ = false;
}

SweepStats DmrgEngine::sweep_single_site() {
  if (site_.empty() || L_.empty() || L_[0].t.n_blocks() == 0) throw ArgError("single-site sweep: prepare() first");
  SweepStats s;
  cur_ = &s;
  steps_.clear();
  const double t0 = now();
  centre_to_site0();
  for (int i = 0; i + 1 < n_; ++i) {
    check_cancel();
    steps_.push_back(step_single_site(i, true));
  }
  for (int i = n_ - 1; i >= 1; --i) {
    check_cancel();
    steps_.push_back(step_single_site(i, false));
  }
  ctx_->sync();
  s.seconds = now() - t0;
  for (const StepStats& st : steps_) {
    ++s.steps;
    s.max_truncation_error = std::max(s.max_truncation_error, st.truncation_error);
    s.max_bond_dim = std::max(s.max_bond_dim, st.bond_dim);
    s.matvecs += st.matvecs;
    s.matvecs_fp64 += st.matvecs_fp64;
    if (!st.converged) ++s.unconverged;
  }
  energy_ = steps_.back().energy;
  s.energy = energy_;
  if (expand_alpha_ > 0 && expand_adaptive_) {
    const double a2 = expand_alpha_ * expand_alpha_;
    if (expand_have_prev_ && std::fabs(energy_ - expand_e_prev_) <= 10.0 * a2 * std::max(1.0, std::fabs(energy_))) {
      expand_alpha_ *= 0.1;
      if (expand_alpha_ < 1e-10) expand_alpha_ = 0.0;
    }
    expand_e_prev_ = energy_;
    expand_have_prev_ = true;
  }
  ++sweeps_;
  cur_ = nullptr;
  return s;
}

double DmrgEngine::apply_site0_operator(const std::vector<double>& op, QN dq) {
  centre_to_site0();
  const QIndex& P = cfg_.mpo[0].phys;
  const int64_t d = P.total;
  if (int64_t(op.size()) != d * d) throw ArgError("site operator: expected a d x d matrix");
  IndexMap pm(P);
  for (int64_t so = 0; so < d; ++so)
    for (int64_t si = 0; si < d; ++si)
      if (op[size_t(so * d + si)] != 0.0 && P.q[pm.sector[so]] != fuse(P.q[pm.sector[si]], dq))
        throw ArgError("site operator: an entry does not carry the stated charge");
  auto shifted = [&](const QIndex& b) {
    std::vector<std::pair<QN, int64_t>> v;
    for (int s = 0; s < b.n(); ++s) v.push_back({fuse(b.q[s], dq), b.dim[s]});
    return QIndex(v);
  };
  std::vector<QIndex> nb(bonds_);
  for (int k = 1; k <= n_; ++k) nb[k] = shifted(bonds_[k]);
  std::vector<Site> ns(n_);
  for (int i = 0; i < n_; ++i) {
    ns[i].t = make_block_table(cfg_.sym, {cfg_.mpo[i].phys, nb[i], nb[i + 1]}, {Dir::Out, Dir::Out, Dir::In}, QN{});
    ns[i].d.reset(ctx_, size_t(std::max<int64_t>(1, ns[i].t.total)));
    launch_fill(ctx_, ns[i].d.get(), 0.0, std::max<int64_t>(1, ns[i].t.total));
  }
  // sites 1.. keep their data: the sector order of a uniformly shifted bond
  // is unchanged, so every block moves to the block of the same sectors
  for (int i = 1; i < n_; ++i) {
    std::vector<CopyJob> cj;
    const Site& o = site_[i];
    for (int64_t b = 0; b < o.t.n_blocks(); ++b) {
      int32_t secs[3] = {o.t.s(b, 0), o.t.s(b, 1), o.t.s(b, 2)};
      const int64_t nbk = ns[i].t.find(secs);
      if (nbk < 0) throw std::logic_error("site operator: shifted block missing");
      const int32_t rows = int32_t(cfg_.mpo[i].phys.dim[secs[0]] * bonds_[i].dim[secs[1]]);
      const int32_t cols = int32_t(bonds_[i + 1].dim[secs[2]]);
      cj.push_back({o.t.offset[b], ns[i].t.offset[nbk], rows, cols, cols, cols, 0, 0});
    }
    launch_copies(ctx_, cj, o.d.get(), o.t.total, ns[i].d.get(), ns[i].t.total);
  }
  // site 0: new row (so, r) = sum_si O[so][si] old row (si, r), one output per
  // (so, right sector) and one term per contributing si (coefficient O)
  {
    const Site& o = site_[0];
    DBuf<double> one(ctx_, 1);
    launch_fill(ctx_, one.get(), 1.0, 1);
    ContractList cl;
    cl.flags = 0;
    for (int64_t b = 0; b < ns[0].t.n_blocks(); ++b) {
      const int s_out = ns[0].t.s(b, 0), l = ns[0].t.s(b, 1), r = ns[0].t.s(b, 2);
      const int64_t dL = nb[0].dim[l], dR = nb[1].dim[r];
      for (int64_t xo = 0; xo < P.dim[s_out]; ++xo) {
        const int64_t so = P.off[s_out] + xo;
        for (int64_t row = 0; row < dL; ++row) {
          cl.begin_out(ns[0].t.offset[b] + (xo * dL + row) * dR, 1, int(dR), int(dR));
          for (int64_t si = 0; si < d; ++si) {
            const double c = op[size_t(so * d + si)];
            if (c == 0.0) continue;
            int32_t secs[3] = {pm.sector[si], l, r};
            const int64_t ob = o.t.find(secs);
            if (ob < 0) continue;
            cl.begin_term(o.t.offset[ob] + (int64_t(pm.within[si]) * dL + row) * dR, int(dR), 1);
            cl.add_src(0, 1, c);
          }
        }
      }
    }
    if (!cl.outs.empty()) run_list(ctx_, cl, 1, o.t.total, ns[0].t.total, one.get(), o.d.get(), ns[0].d.get());
  }
  DBuf<double> nrm(ctx_, 1);
  dot_many(ctx_, ns[0].d.get(), ns[0].t.total, 1, ns[0].d.get(), ns[0].t.total, nrm.get());
  double norm2 = 0;
  nrm.download(&norm2, 1);
  ctx_->sync();
  if (!(norm2 > 0) || !std::isfinite(norm2)) throw ArgError("site operator: the result is zero or not finite");
  scale(ctx_, ns[0].d.get(), 1.0 / std::sqrt(norm2), ns[0].t.total);
  bonds_ = std::move(nb);
  site_ = std::move(ns);
  L_.clear();
  R_.clear();
  L_.resize(n_ + 1);
  R_.resize(n_ + 1);
  L_[0] = boundary_env(bonds_[0], cfg_.mpo[0].wl);
  R_[n_] = boundary_env(bonds_[n_], cfg_.mpo[n_ - 1].wr);
  for (int i = n_ - 1; i >= 1; --i) grow_right(i);
  ctx_->sync();
  centre_ = 0;
  return norm2;
}

double mps_overlap(const DmrgEngine& bra, const DmrgEngine& ket) {
  Context* ctx = ket.context();
  const int n = ket.n_sites();
  if (bra.n_sites() != n || bra.sym() != ket.sym()) throw ArgError("overlap: the two states differ in length or symmetry");
  if (bra.bonds().empty() || ket.bonds().empty()) throw ArgError("overlap: set_state first");
  for (int i = 0; i < n; ++i)
    if (!same_qindex(bra.phys(i), ket.phys(i))) throw ArgError("overlap: physical legs differ at site " + std::to_string(i));
  // E on one bond: one block per charge present in both bond bases
  struct EBlock {
    int sb, sk;  // sector in the bra / ket bond
    int64_t off, db, dk;
  };
  auto layout = [&](const QIndex& qb, const QIndex& qk, std::vector<EBlock>& blocks) {
    blocks.clear();
    int64_t tot = 0;
    for (int a = 0; a < qb.n(); ++a) {
      const int b = qk.find(qb.q[a]);
      if (b < 0) continue;
      blocks.push_back({a, b, tot, qb.dim[a], qk.dim[b]});
      tot += qb.dim[a] * qk.dim[b];
    }
    return tot;
  };
  std::vector<EBlock> eb, nb;
  int64_t etot = layout(bra.bonds()[0], ket.bonds()[0], eb);
  if (etot != 1) throw ArgError("overlap: the left boundary bonds differ");
  DBuf<double> E(ctx, 1);
  launch_fill(ctx, E.get(), 1.0, 1);
  for (int i = 0; i < n; ++i) {
    const QIndex& P = ket.phys(i);
    const QIndex& bbL = bra.bonds()[i];
    const QIndex& bbR = bra.bonds()[i + 1];
    const QIndex& bkL = ket.bonds()[i];
    const QIndex& bkR = ket.bonds()[i + 1];
    const BlockTable& tb = bra.site_table(i);
    const BlockTable& tk = ket.site_table(i);
    const int64_t ntot = layout(bbR, bkR, nb);
    std::map<int, const EBlock*> e_of_ket;  // ket bond sector -> E block
    for (const EBlock& x : eb) e_of_ket[x.sk] = &x;
    std::map<int, const EBlock*> n_of_ket;
    for (const EBlock& x : nb) n_of_ket[x.sk] = &x;
    // stage 1: T = E[q] A_ket[s, x](q, q') for every ket block and x
    struct TSlice {
      int64_t off, db, dkr;
      int s, x, bsec_l, ksec_r;
    };
    std::vector<TSlice> ts;
    int64_t ttot = 0;
    ContractList c1;
    c1.flags = 0;
    for (int64_t b = 0; b < tk.n_blocks(); ++b) {
      const int s = tk.s(b, 0), l = tk.s(b, 1), r = tk.s(b, 2);
      auto ei = e_of_ket.find(l);
      if (ei == e_of_ket.end() || n_of_ket.find(r) == n_of_ket.end()) continue;
      const EBlock& e = *ei->second;
      const int64_t dkl = bkL.dim[l], dkr = bkR.dim[r];
      for (int x = 0; x < P.dim[s]; ++x) {
        ts.push_back({ttot, e.db, dkr, s, x, e.sb, r});
        c1.begin_out(ttot, int(e.db), int(dkr), int(dkr));
        c1.begin_term(tk.offset[b] + int64_t(x) * dkl * dkr, int(dkr), int(dkl));
        c1.add_src(e.off, int(dkl), 1.0);
        ttot += e.db * dkr;
      }
    }
    DBuf<double> T(ctx, size_t(std::max<int64_t>(1, ttot)));
    DBuf<double> En(ctx, size_t(std::max<int64_t>(1, ntot)));
    launch_fill(ctx, En.get(), 0.0, std::max<int64_t>(1, ntot));
    if (!c1.outs.empty()) run_list(ctx, c1, etot, tk.total, ttot, E.get(), ket.site_device(i), T.get());
    // stage 2: E'[q'] = sum over (s, x, q) of A_bra[s, x](q, q')^T T
    ContractList c2;
    c2.flags = kTransA;
    for (const EBlock& o : nb) {
      bool opened = false;
      for (const TSlice& t : ts) {
        if (t.ksec_r != o.sk) continue;
        int32_t secs[3] = {t.s, t.bsec_l, o.sb};
        const int64_t bb = tb.find(secs);
        if (bb < 0) continue;
        if (!opened) {
          c2.begin_out(o.off, int(o.db), int(o.dk), int(o.dk));
          opened = true;
        }
        const int64_t dbl = bbL.dim[t.bsec_l], dbr = bbR.dim[o.sb];
        c2.begin_term(t.off, int(t.dkr), int(dbl));
        c2.add_src(tb.offset[bb] + int64_t(t.x) * dbl * dbr, int(dbr), 1.0);
      }
    }
    if (!c2.outs.empty()) run_list(ctx, c2, tb.total, ttot, ntot, bra.site_device(i), T.get(), En.get());
    E = std::move(En);
    eb = nb;
    etot = ntot;
  }
  if (etot != 1) throw ArgError("overlap: the states end in different total charges");
  double v = 0;
  E.download(&v, 1);
  ctx->sync();
  return v;
}

TdvpStats DmrgEngine::tdvp_step(double tau, int64_t max_bond, int krylov_dim, double krylov_tol,
                                double adaptive_tol) {
  if (site_.empty() || L_.empty() || L_[0].t.n_blocks() == 0) throw ArgError("tdvp: prepare() first");
  if (!std::isfinite(tau)) throw ArgError("tdvp: time step is not finite");
  centre_to_site0();
  if (max_bond <= 0) max_bond = cfg_.bond_schedule.back();
  TdvpStats ts;
  const double t0 = now();
  const double floor = std::max(cfg_.svd_cutoff * cfg_.svd_cutoff, kLambdaZero);
  // two-site update of bond (i, i+1): evolve, split, grow the environment
  // behind the centre; the last bond of the step also gives the energy
  auto two_site = [&](int i, bool lr, bool energy) {
    const double t0 = now();
    HeffPlan plan(ctx_, Prec::FP64, L_[i].t, L_[i].d.get(), R_[i + 2].t, R_[i + 2].d.get(), cfg_.mpo[i],
                  cfg_.mpo[i + 1]);
    const double t1 = now();
    const BlockTable& psi = plan.psi();
    const size_t len = size_t(std::max<int64_t>(1, psi.total));
    DBuf<double> x(ctx_, len), y(ctx_, len);
    contract_two_site(i, psi, x.get());
    const ExpKrylovResult r =
        exp_krylov(ctx_, plan, x.get(), -0.5 * tau, krylov_dim, krylov_tol, y.get(), adaptive_tol);
    ts.matvecs += r.matvecs;
    if (!(r.norm_out > 0) || !(r.norm_in > 0)) throw ArgError("tdvp: the two-site tensor vanished");
    ts.log_norm += std::log(r.norm_out) - std::log(r.norm_in);
    if (energy) {
      DBuf<double> hy(ctx_, len), d(ctx_, 2);
      plan.apply_fp64(y.get(), hy.get());
      ++ts.matvecs;
      dot_many(ctx_, y.get(), psi.total, 1, hy.get(), psi.total, d.get());
      dot_many(ctx_, y.get(), psi.total, 1, y.get(), psi.total, d.get() + 1);
      double h[2];
      d.download(h, 2);
      ctx_->sync();
      ts.energy = h[0] / h[1];
    }
    const double t2 = now();
    StepStats st;
    split_two_site(i, lr, psi, y.get(), max_bond, cfg_.svd_cutoff, floor, &st);
    ts.max_truncation_error = std::max(ts.max_truncation_error, st.truncation_error);
    ts.max_bond_dim = std::max(ts.max_bond_dim, st.bond_dim);
    const double t3 = now();
    if (lr)
      grow_left(i);
    else
      grow_right(i + 1);
    ctx_->sync();
    ts.t_plan += t1 - t0;
    ts.t_exp += t2 - t1;
    ts.t_svd += t3 - t2;
    ts.t_env += now() - t3;
  };
  for (int i = 0; i + 1 < n_; ++i) {
    two_site(i, true, false);
    if (i + 2 < n_) ts.log_norm += evolve_one_site(i + 1, 0.5 * tau, krylov_dim, krylov_tol, adaptive_tol, &ts);
  }
  for (int i = n_ - 2; i >= 0; --i) {
    two_site(i, false, i == 0);
    if (i > 0) ts.log_norm += evolve_one_site(i, 0.5 * tau, krylov_dim, krylov_tol, adaptive_tol, &ts);
  }
  ctx_->sync();
  energy_ = ts.energy;
  ts.seconds = now() - t0;
  return ts;
}

void DmrgEngine::contract_two_site(int i, const BlockTable& psi, double* x) {
  const Site& A = site_[i];
  const Site& B = site_[i + 1];
  const QIndex& Pi = cfg_.mpo[i].phys;
  const QIndex& Pj = cfg_.mpo[i + 1].phys;
  const QIndex& BL = bonds_[i];
  const QIndex& BM = bonds_[i + 1];
  const QIndex& BR = bonds_[i + 2];
  ContractList cl;
  cl.flags = 0;
  for (int64_t b = 0; b < psi.n_blocks(); ++b) {
    const int si = psi.s(b, 0), sj = psi.s(b, 1), ls = psi.s(b, 2), rs = psi.s(b, 3);
    const int64_t dL = BL.dim[ls], dR = BR.dim[rs], dsj = Pj.dim[sj];
    const int ms = BM.find(fuse(BL.q[ls], Pi.q[si]));
    int64_t ab = -1, bb = -1;
    if (ms >= 0) {
      int32_t as[3] = {si, ls, ms}, bs[3] = {sj, ms, rs};
      ab = A.t.find(as);
      bb = B.t.find(bs);
    }
    for (int64_t xi = 0; xi < Pi.dim[si]; ++xi)
      for (int64_t xj = 0; xj < dsj; ++xj) {
        cl.begin_out(psi.offset[b] + (xi * dsj + xj) * dL * dR, int(dL), int(dR), int(dR));
        if (ab < 0 || bb < 0) continue;
        const int64_t dm = BM.dim[ms];
        cl.begin_term(B.t.offset[bb] + xj * dm * dR, int(dR), int(dm));
        cl.add_src(A.t.offset[ab] + xi * dL * dm, int(dm), 1.0);
      }
  }
  run_list(ctx_, cl, A.t.total, B.t.total, psi.total, A.d.get(), B.d.get(), x);
}

void DmrgEngine::split_two_site(int i, bool lr, const BlockTable& psi, const double* x, int64_t max_bond,
                                double cutoff, double lambda_floor, StepStats* st) {
  const QIndex& Pi = cfg_.mpo[i].phys;
  const QIndex& Pj = cfg_.mpo[i + 1].phys;
  const QIndex& BL = bonds_[i];
  const QIndex& BR = bonds_[i + 2];
  // Middle charges of the (s_i, a_L) x (s_j, a_R) matricisation.  Row parts in
  // (s_i sector, x_i) order, column parts in (s_j sector, x_j) order: the
  // offsets of fuse_legs (block_sparse.rs:474-518) for a fixed fused charge.
  struct Part {
    int ssec, x, bsec;
    int64_t off, d;
  };
  struct Mid {
    QN q;
    std::vector<Part> rows, cols;
    int64_t nrow = 0, ncol = 0, n = 0, g_off = 0, w_off = 0;
    int64_t keep = 0;
  };
  std::map<QN, Mid> mm;
  for (int s = 0; s < Pi.n(); ++s)
    for (int x = 0; x < Pi.dim[s]; ++x)
      for (int l = 0; l < BL.n(); ++l) {
        Mid& m = mm[fuse(Pi.q[s], BL.q[l])];
        m.rows.push_back({s, x, l, m.nrow, BL.dim[l]});
        m.nrow += BL.dim[l];
      }
  for (int s = 0; s < Pj.n(); ++s)
    for (int x = 0; x < Pj.dim[s]; ++x)
      for (int r = 0; r < BR.n(); ++r) {
        Mid& m = mm[fuse(BR.q[r], dual(Pj.q[s]))];
        m.cols.push_back({s, x, r, m.ncol, BR.dim[r]});
        m.ncol += BR.dim[r];
      }
  std::vector<Mid> mids;
  for (auto& [q, m] : mm)
    if (m.nrow > 0 && m.ncol > 0) {
      m.q = q;
      mids.push_back(std::move(m));
    }
  // dense block of every middle charge: X = M (left-to-right) or M^T
  // (right-to-left), so the kept vectors are always the left ones of X
  int64_t xtot = 0;
  std::vector<int64_t> x_off;
  for (Mid& m : mids) {
    m.n = lr ? m.nrow : m.ncol;
    if (m.n > kSymEigMaxN) throw ArgError("SVD block larger than the eigensolver limit");
    x_off.push_back(xtot);
    xtot += m.nrow * m.ncol;
  }
  auto slice = [&](int ssi, int xi, int ls, int ssj, int xj, int rs) {
    int32_t secs[4] = {ssi, ssj, ls, rs};
    const int64_t b = psi.find(secs);
    if (b < 0) throw std::logic_error("split: missing psi block");
    return psi.offset[b] + (int64_t(xi) * Pj.dim[ssj] + xj) * BL.dim[ls] * BR.dim[rs];
  };
  DBuf<double> X(ctx_, size_t(std::max<int64_t>(1, xtot)));
  {
    std::vector<CopyJob> xc;
    for (size_t k = 0; k < mids.size(); ++k) {
      const Mid& m = mids[k];
      for (const Part& rp : m.rows)
        for (const Part& cp : m.cols) {
          const int64_t src = slice(rp.ssec, rp.x, rp.bsec, cp.ssec, cp.x, cp.bsec);
          if (lr)
            xc.push_back({src, x_off[k] + rp.off * m.ncol + cp.off, int32_t(rp.d), int32_t(cp.d), int32_t(cp.d),
                          int32_t(m.ncol), 0, 0});
          else
            xc.push_back({src, x_off[k] + cp.off * m.nrow + rp.off, int32_t(cp.d), int32_t(rp.d), int32_t(cp.d),
                          int32_t(m.nrow), 1, 0});
        }
    }
    launch_copies(ctx_, xc, x, psi.total, X.get(), xtot);
  }
  // squared singular values of every block (descending), two-level Gram,
  // the global selection of truncate_svd (truncation/mod.rs:72-82) with the
  // degeneracy and numerical-zero rules, then the kept vectors
  std::vector<GramBlock> gb;
  for (size_t k = 0; k < mids.size(); ++k)
    gb.push_back({x_off[k], int32_t(mids[k].n), int32_t(lr ? mids[k].ncol : mids[k].nrow)});
  TruncRule rule;
  rule.max_bond = max_bond;
  rule.cutoff = cutoff;
  rule.min_bond = cfg_.min_bond_dim;
  rule.degeneracy_tol = cfg_.degeneracy_tol;
  rule.lambda_floor = lambda_floor;
  BlockSvd bs(ctx_);
  const TruncSelection& sel = bs.run(gb, X.get(), xtot, rule);
  if (!(sel.lambda_max > 0) || !std::isfinite(sel.lambda_max))
    throw ArgError("split: the two-site tensor is zero or not finite (zero-norm MPS?)");
  for (size_t k = 0; k < mids.size(); ++k) {
    mids[k].g_off = bs.u_offset(k);
    mids[k].keep = sel.keep[k];
  }
  const int64_t gtot = bs.u_size();
  st->truncation_error = sel.rel_error();
  const double renorm = sel.sum_kept > 0 ? 1.0 / std::sqrt(sel.sum_kept) : 1.0;
  const int ph_write = ctx_->timer.begin(ctx_->stream);
  const double* E = bs.u();
  // new bond
  std::vector<std::pair<QN, int64_t>> nb;
  std::map<QN, int> mid_of;
  for (int k = 0; k < int(mids.size()); ++k)
    if (mids[k].keep > 0) {
      nb.push_back({mids[k].q, mids[k].keep});
      mid_of[mids[k].q] = k;
    }
  QIndex newB(nb);
  st->bond_dim = newB.total;
  st->dim = psi.total;
  for (const Mid& m : mids) {
    st->svd_nmax = std::max(st->svd_nmax, m.n);
    st->svd_n3 += double(m.n) * double(m.n) * double(m.n);
  }
  Site Ai, Aj;
  Ai.t = make_block_table(cfg_.sym, {Pi, BL, newB}, {Dir::Out, Dir::Out, Dir::In}, QN{});
  Aj.t = make_block_table(cfg_.sym, {Pj, newB, BR}, {Dir::Out, Dir::Out, Dir::In}, QN{});
  Ai.d.reset(ctx_, size_t(std::max<int64_t>(1, Ai.t.total)));
  Aj.d.reset(ctx_, size_t(std::max<int64_t>(1, Aj.t.total)));
  // locate the row/column part of a site block
  auto row_part = [&](const Mid& m, int s, int x) -> const Part& {
    for (const Part& p : m.rows)
      if (p.ssec == s && p.x == x) return p;
    throw std::logic_error("split: row part not found");
  };
  auto col_part = [&](const Mid& m, int s, int x) -> const Part& {
    for (const Part& p : m.cols)
      if (p.ssec == s && p.x == x) return p;
    throw std::logic_error("split: column part not found");
  };
  std::vector<CopyJob> copies;
  ContractList wl;
  if (lr) {
    // A_i = U (left-canonical): copy the kept eigenvector columns
    for (int64_t b = 0; b < Ai.t.n_blocks(); ++b) {
      const int s = Ai.t.s(b, 0), l = Ai.t.s(b, 1), ms = Ai.t.s(b, 2);
      const Mid& m = mids[mid_of.at(newB.q[ms])];
      const int64_t dL = BL.dim[l], k = m.keep;
      for (int x = 0; x < Pi.dim[s]; ++x) {
        const Part& p = row_part(m, s, x);
        copies.push_back({m.g_off + p.off * m.n, Ai.t.offset[b] + x * dL * k, int32_t(dL), int32_t(k), int32_t(m.n),
                          int32_t(k), 0, 0});
      }
    }
    // A_j = U^T theta / |kept| (the new centre)
    wl.flags = kTransA;
    for (int64_t b = 0; b < Aj.t.n_blocks(); ++b) {
      const int s = Aj.t.s(b, 0), ms = Aj.t.s(b, 1), r = Aj.t.s(b, 2);
      const Mid& m = mids[mid_of.at(newB.q[ms])];
      const int64_t dR = BR.dim[r], k = m.keep;
      for (int x = 0; x < Pj.dim[s]; ++x) {
        wl.begin_out(Aj.t.offset[b] + x * k * dR, int(k), int(dR), int(dR));
        for (const Part& rp : m.rows) {
          wl.begin_term(slice(rp.ssec, rp.x, rp.bsec, s, x, r), int(dR), int(rp.d));
          wl.add_src(m.g_off + rp.off * m.n, int(m.n), renorm);
        }
      }
    }
    launch_copies(ctx_, copies, E, gtot, Ai.d.get(), Ai.t.total);
    run_list(ctx_, wl, gtot, psi.total, Aj.t.total, E, x, Aj.d.get());
  } else {
    // A_j = V^T (right-canonical)
    for (int64_t b = 0; b < Aj.t.n_blocks(); ++b) {
      const int s = Aj.t.s(b, 0), ms = Aj.t.s(b, 1), r = Aj.t.s(b, 2);
      const Mid& m = mids[mid_of.at(newB.q[ms])];
      const int64_t dR = BR.dim[r], k = m.keep;
      for (int x = 0; x < Pj.dim[s]; ++x) {
        const Part& p = col_part(m, s, x);
        copies.push_back({m.g_off + p.off * m.n, Aj.t.offset[b] + x * k * dR, int32_t(k), int32_t(dR), int32_t(m.n),
                          int32_t(dR), 1, 0});
      }
    }
    // A_i = theta V / |kept| (the new centre)
    wl.flags = 0;
    for (int64_t b = 0; b < Ai.t.n_blocks(); ++b) {
      const int s = Ai.t.s(b, 0), l = Ai.t.s(b, 1), ms = Ai.t.s(b, 2);
      const Mid& m = mids[mid_of.at(newB.q[ms])];
      const int64_t dL = BL.dim[l], k = m.keep;
      for (int x = 0; x < Pi.dim[s]; ++x) {
        wl.begin_out(Ai.t.offset[b] + x * dL * k, int(dL), int(k), int(k));
        for (const Part& cp : m.cols) {
          wl.begin_term(m.g_off + cp.off * m.n, int(m.n), int(cp.d));
          wl.add_src(slice(s, x, l, cp.ssec, cp.x, cp.bsec), int(cp.d), renorm);
        }
      }
    }
    launch_copies(ctx_, copies, E, gtot, Aj.d.get(), Aj.t.total);
    run_list(ctx_, wl, psi.total, gtot, Ai.t.total, x, E, Ai.d.get());
  }
  ctx_->timer.end(ph_write, ctx_->stream, "phase_svd_writeback", 0.0, 0.0, 1);
  centre_ = lr ? i + 1 : i;
  bonds_[i + 1] = newB;
  site_[i] = std::move(Ai);
  site_[i + 1] = std::move(Aj);
}

void DmrgEngine::grow_left(int i) {
  // L'[a'_R, w_r, a_R] = sum A[s'](a'_L, a'_R) L[a'_L, w_l, a_L] W[s, s', w_l, w_r] A[s](a_L, a_R)
  // (environments/mod.rs:225-355 in block-sparse form)
  const Env& Lo = L_[i];
  const Site& A = site_[i];
  const MpoSite& W = cfg_.mpo[i];
  const QIndex& P = W.phys;
  const QIndex& BL = bonds_[i];
  const QIndex& BR = bonds_[i + 1];
  Env out;
  out.t = make_block_table(cfg_.sym, {BR, W.wr, BR}, {Dir::In, Dir::In, Dir::Out}, QN{});
  out.d.reset(ctx_, size_t(std::max<int64_t>(1, out.t.total)));
  IndexMap PM(P), WLm(W.wl), WRm(W.wr);
  // W entries grouped by (s_out, w_r)
  std::vector<std::vector<std::tuple<int, int, double>>> src(size_t(P.total * W.wr.total));
  for (int64_t s = 0; s < P.total; ++s)
    for (int64_t so = 0; so < P.total; ++so)
      for (int64_t a = 0; a < W.wl.total; ++a)
        for (int64_t b = 0; b < W.wr.total; ++b) {
          const double v = W.at(s, so, a, b);
          if (v != 0.0) src[size_t(so * W.wr.total + b)].push_back({int(s), int(a), v});
        }
  // T1 registry keyed by (site block, x_s, w_l)
  std::map<std::tuple<int64_t, int, int>, int64_t> t1_off;
  struct T1 { int64_t off, b_off, a_off; int m, n, k, lda; };
  std::vector<T1> t1s;
  int64_t t1_total = 0;
  ContractList s2;
  s2.flags = kTransA | kTransC;
  for (int64_t ob = 0; ob < out.t.n_blocks(); ++ob) {
    const int rps = out.t.s(ob, 0), swr = out.t.s(ob, 1), rs = out.t.s(ob, 2);
    const int64_t dRp = BR.dim[rps], dR = BR.dim[rs], dwr = W.wr.dim[swr];
    for (int64_t iwr = 0; iwr < dwr; ++iwr) {
      const int64_t wr = W.wr.off[swr] + iwr;
      s2.begin_out(out.t.offset[ob] + iwr * dR, int(dR), int(dRp), int(dwr * dR));
      for (int sps = 0; sps < P.n(); ++sps)
        for (int xps = 0; xps < P.dim[sps]; ++xps) {
          const QN qlp = fuse(BR.q[rps], dual(P.q[sps]));
          const int lps = BL.find(qlp);
          if (lps < 0) continue;
          int32_t bsec[3] = {sps, lps, rps};
          const int64_t bb = A.t.find(bsec);
          if (bb < 0) continue;
          const int64_t dLp = BL.dim[lps];
          s2.begin_term(A.t.offset[bb] + xps * dLp * dRp, int(dRp), int(dLp));
          const int64_t so = P.off[sps] + xps;
          for (auto& [s, wlv, coef] : src[size_t(so * W.wr.total + wr)]) {
            const int ssec = PM.sector[s], xs = PM.within[s];
            const int swl = WLm.sector[wlv], iwl = WLm.within[wlv];
            const int ls = BL.find(fuse(qlp, W.wl.q[swl]));
            if (ls < 0) continue;
            int32_t ksec[3] = {ssec, ls, rs};
            const int64_t kb = A.t.find(ksec);
            if (kb < 0) continue;
            int32_t lsec[3] = {lps, swl, ls};
            const int64_t lb = Lo.t.find(lsec);
            if (lb < 0) continue;
            auto key = std::make_tuple(kb, xs, int(wlv));
            auto it = t1_off.find(key);
            if (it == t1_off.end()) {
              const int64_t dL = BL.dim[ls];
              T1 t{t1_total, A.t.offset[kb] + xs * dL * dR, Lo.t.offset[lb] + iwl * dL, int(dLp), int(dR), int(dL),
                   int(W.wl.dim[swl] * dL)};
              t1s.push_back(t);
              it = t1_off.emplace(key, t1_total).first;
              t1_total += dLp * dR;
            }
            s2.add_src(it->second, int(dR), coef);
          }
          s2.drop_empty_term();
        }
    }
  }
  ContractList s1;
  s1.flags = 0;
  for (const T1& t : t1s) {
    s1.begin_out(t.off, t.m, t.n, t.n);
    s1.begin_term(t.b_off, t.n, t.k);
    s1.add_src(t.a_off, t.lda, 1.0);
  }
  DBuf<double> T(ctx_, size_t(std::max<int64_t>(1, t1_total)));
  run_list(ctx_, s1, Lo.t.total, A.t.total, t1_total, Lo.d.get(), A.d.get(), T.get());
  run_list(ctx_, s2, t1_total, A.t.total, out.t.total, T.get(), A.d.get(), out.d.get());
  L_[i + 1] = std::move(out);
}

void DmrgEngine::grow_right(int i) {
  // R'[a'_L, w_l, a_L] = sum A[s'](a'_L, a'_R) R[a'_R, w_r, a_R] W[s, s', w_l, w_r] A[s](a_L, a_R)
  // (environments/mod.rs:361-483 in block-sparse form)
  const Env& Ro = R_[i + 1];
  const Site& A = site_[i];
  const MpoSite& W = cfg_.mpo[i];
  const QIndex& P = W.phys;
  const QIndex& BL = bonds_[i];
  const QIndex& BR = bonds_[i + 1];
  Env out;
  out.t = make_block_table(cfg_.sym, {BL, W.wl, BL}, {Dir::In, Dir::In, Dir::Out}, QN{});
  out.d.reset(ctx_, size_t(std::max<int64_t>(1, out.t.total)));
  IndexMap PM(P), WLm(W.wl), WRm(W.wr);
  // W entries grouped by (s_out, w_l)
  std::vector<std::vector<std::tuple<int, int, double>>> src(size_t(P.total * W.wl.total));
  for (int64_t s = 0; s < P.total; ++s)
    for (int64_t so = 0; so < P.total; ++so)
      for (int64_t a = 0; a < W.wl.total; ++a)
        for (int64_t b = 0; b < W.wr.total; ++b) {
          const double v = W.at(s, so, a, b);
          if (v != 0.0) src[size_t(so * W.wl.total + a)].push_back({int(s), int(b), v});
        }
  std::map<std::tuple<int64_t, int, int>, int64_t> t1_off;
  struct T1 { int64_t off, b_off, a_off; int m, n, k, lda, ldb; };
  std::vector<T1> t1s;
  int64_t t1_total = 0;
  ContractList s2;
  s2.flags = kTransA | kTransB | kTransC;
  for (int64_t ob = 0; ob < out.t.n_blocks(); ++ob) {
    const int lps = out.t.s(ob, 0), swl = out.t.s(ob, 1), ls = out.t.s(ob, 2);
    const int64_t dLp = BL.dim[lps], dL = BL.dim[ls], dwl = W.wl.dim[swl];
    for (int64_t iwl = 0; iwl < dwl; ++iwl) {
      const int64_t wlv = W.wl.off[swl] + iwl;
      s2.begin_out(out.t.offset[ob] + iwl * dL, int(dL), int(dLp), int(dwl * dL));
      for (int sps = 0; sps < P.n(); ++sps)
        for (int xps = 0; xps < P.dim[sps]; ++xps) {
          const QN qrp = fuse(BL.q[lps], P.q[sps]);
          const int rps = BR.find(qrp);
          if (rps < 0) continue;
          int32_t bsec[3] = {sps, lps, rps};
          const int64_t bb = A.t.find(bsec);
          if (bb < 0) continue;
          const int64_t dRp = BR.dim[rps];
          s2.begin_term(A.t.offset[bb] + xps * dLp * dRp, int(dRp), int(dRp));
          const int64_t so = P.off[sps] + xps;
          for (auto& [s, wrv, coef] : src[size_t(so * W.wl.total + wlv)]) {
            const int ssec = PM.sector[s], xs = PM.within[s];
            const int swr = WRm.sector[wrv], iwr = WRm.within[wrv];
            const int rs = BR.find(fuse(qrp, W.wr.q[swr]));
            if (rs < 0) continue;
            int32_t ksec[3] = {ssec, ls, rs};
            const int64_t kb = A.t.find(ksec);
            if (kb < 0) continue;
            int32_t rsec[3] = {rps, swr, rs};
            const int64_t rb = Ro.t.find(rsec);
            if (rb < 0) continue;
            auto key = std::make_tuple(kb, xs, int(wrv));
            auto it = t1_off.find(key);
            if (it == t1_off.end()) {
              const int64_t dR = BR.dim[rs];
              T1 t{t1_total, A.t.offset[kb] + xs * dL * dR, Ro.t.offset[rb] + iwr * dR, int(dRp), int(dL), int(dR),
                   int(W.wr.dim[swr] * dR), int(dR)};
              t1s.push_back(t);
              it = t1_off.emplace(key, t1_total).first;
              t1_total += dRp * dL;
            }
            s2.add_src(it->second, int(dL), coef);
          }
          s2.drop_empty_term();
        }
    }
  }
  ContractList s1;
  s1.flags = kTransB;
  for (const T1& t : t1s) {
    s1.begin_out(t.off, t.m, t.n, t.n);
    s1.begin_term(t.b_off, t.ldb, t.k);
    s1.add_src(t.a_off, t.lda, 1.0);
  }
  DBuf<double> T(ctx_, size_t(std::max<int64_t>(1, t1_total)));
  run_list(ctx_, s1, Ro.t.total, A.t.total, t1_total, Ro.d.get(), A.d.get(), T.get());
  run_list(ctx_, s2, t1_total, A.t.total, out.t.total, T.get(), A.d.get(), out.d.get());
  R_[i] = std::move(out);
}

}  // namespace tk
