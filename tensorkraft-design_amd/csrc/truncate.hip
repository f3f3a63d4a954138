// SVD truncation rule, bond schedule and the truncated SVDs built on the
// two-level Gram eigen-decomposition (declarations and method in truncate.hpp).
#include "truncate.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>

#include "backend.hpp"
#include "kernels.hpp"
#include "tdeig.hpp"

namespace tk {

TruncSelection select_truncation(const std::vector<std::vector<double>>& lambda, const TruncRule& rule) {
  struct Cand {
    double l;
    int blk, idx;
  };
  std::vector<Cand> all;
  for (int b = 0; b < int(lambda.size()); ++b)
    for (int j = 0; j < int(lambda[b].size()); ++j) all.push_back({std::max(0.0, lambda[b][j]), b, j});
  std::stable_sort(all.begin(), all.end(), [](const Cand& a, const Cand& b) {
    if (a.l != b.l) return a.l > b.l;
    if (a.blk != b.blk) return a.blk < b.blk;
    return a.idx < b.idx;
  });
  TruncSelection s;
  s.keep.assign(lambda.size(), 0);
  if (all.empty()) return s;
  const double lmax = all[0].l;
  s.lambda_max = lmax;
  // sigma >= cutoff * sigma_max  <=>  sigma^2 >= cutoff^2 sigma_max^2
  const double thr = std::max(rule.cutoff * rule.cutoff, rule.lambda_floor) * lmax;
  size_t keep = 0;
  while (keep < all.size() && int64_t(keep) < rule.max_bond && all[keep].l >= thr) ++keep;
  if (rule.degeneracy_tol > 0 && keep < all.size() && keep > 0) {
    const double smax = std::sqrt(lmax);
    size_t k2 = keep;
    while (k2 > 0 && std::sqrt(all[k2 - 1].l) - std::sqrt(all[k2].l) <= rule.degeneracy_tol * smax) --k2;
    if (k2 > 0) keep = k2;
  }
  keep = std::max<size_t>(keep, size_t(std::min<int64_t>(std::max<int64_t>(rule.min_bond, 0), int64_t(all.size()))));
  for (size_t k = 0; k < all.size(); ++k) {
    s.sum_all += all[k].l;
    if (k < keep) {
      s.sum_kept += all[k].l;
      ++s.keep[all[k].blk];
    }
  }
  s.kept = int64_t(keep);
  s.discarded = int64_t(all.size() - keep);
  return s;
}

namespace {
// f64::powi for a non-negative exponent: square-and-multiply from the low bit
double powi(double a, int64_t b) {
  double r = 1.0;
  while (true) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}
// `x.round() as usize`: half away from zero, NaN and negatives to 0, saturating
int64_t round_to_count(double x) {
  if (!(x > 0)) return 0;
  const double r = std::round(x);
  if (r >= 9.2e18) return std::numeric_limits<int64_t>::max();
  return int64_t(r);
}
}  // namespace

BondSchedule BondSchedule::warmup(int64_t d_init, int64_t d_max, int64_t n_warmup_sweeps) {
  if (d_init < 0 || d_max < 0 || n_warmup_sweeps < 0) throw ArgError("bond schedule: negative argument");
  if (n_warmup_sweeps == 0) return BondSchedule{{d_max}};
  const double ratio = std::pow(double(d_max) / double(d_init), 1.0 / double(n_warmup_sweeps));
  BondSchedule s;
  for (int64_t i = 0; i <= n_warmup_sweeps; ++i) {
    const int64_t d = round_to_count(double(d_init) * powi(ratio, i));
    s.dims.push_back(std::max<int64_t>(std::min(d, d_max), 1));
  }
  return s;
}

int64_t BondSchedule::at_sweep(int64_t sweep_index) const {
  if (sweep_index >= 0 && sweep_index < int64_t(dims.size())) return dims[size_t(sweep_index)];
  return dims.empty() ? 1 : dims.back();
}

const TruncSelection& BlockSvd::run(const std::vector<GramBlock>& blocks, const double* X, int64_t x_n,
                                    const TruncRule& rule) {
  const double lam_min_rel = std::max(rule.cutoff * rule.cutoff, rule.lambda_floor);
  const int ph_values = ctx_->timer.begin(ctx_->stream);
  lambda_ = gs_.values(blocks, X, x_n, lam_min_rel, rule.max_bond);
  ctx_->timer.end(ph_values, ctx_->stream, "phase_svd_values", 0.0, 0.0, 1);
  // every singular value of a block is a candidate; those no Gram level
  // resolved count as zero
  for (size_t b = 0; b < blocks.size(); ++b) {
    const size_t rank = size_t(std::min(blocks[b].n, blocks[b].p));
    if (lambda_[b].size() > rank) lambda_[b].resize(rank);
    while (lambda_[b].size() < rank) lambda_[b].push_back(0.0);
  }
  sel_ = select_truncation(lambda_, rule);
  // a zero (or non-finite) matrix gets no vectors; the caller handles it
  if (!blocks.empty() && sel_.lambda_max > 0 && std::isfinite(sel_.lambda_max)) {
    const int ph_vec = ctx_->timer.begin(ctx_->stream);
    gs_.vectors(sel_.keep);
    ctx_->timer.end(ph_vec, ctx_->stream, "phase_svd_vectors", 0.0, 0.0, 1);
  }
  return sel_;
}

namespace {
// W (p x r, row-major) = C^T diag(inv): W[j*r + i] = C[i*p + j] * inv[i]
__global__ void scaled_transpose_kernel(const double* __restrict__ C, const double* __restrict__ inv, int64_t r,
                                        int64_t p, double* __restrict__ W) {
  const int64_t n = r * p;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t j = e / r, i = e - j * r;
    W[e] = C[i * p + j] * inv[i];
  }
}
}  // namespace

DenseSvd dense_svd_truncate(Context* ctx, int64_t m, int64_t n, const double* A, const TruncRule& rule, double* U,
                            double* Vt) {
  if (m <= 0 || n <= 0) throw ArgError("svd: empty matrix");
  const bool rows_side = m <= n;  // X = A and G = A A^T, else X = A^T and G = A^T A
  const int64_t g = rows_side ? m : n, p = rows_side ? n : m;
  if (g > kSymEigMaxN) throw ArgError("svd: the smaller dimension exceeds 2048");
  const int64_t mn = m * n;
  DBuf<double> Xt;
  const double* X = A;
  if (!rows_side) {
    Xt.reset(ctx, size_t(mn));
    launch_copies(ctx, {CopyJob{0, 0, int32_t(n), int32_t(m), int32_t(n), int32_t(m), 1, 0}}, A, mn, Xt.get(), mn);
    X = Xt.get();
  }
  DenseSvd out;
  BlockSvd bs(ctx);
  out.sel = bs.run({GramBlock{0, int32_t(g), int32_t(p)}}, X, mn, rule);
  const int64_t r = out.sel.kept;
  out.rank = r;
  for (int64_t j = 0; j < r; ++j) out.s.push_back(std::sqrt(std::max(0.0, bs.lambda()[0][size_t(j)])));
  if (!std::isfinite(out.sel.lambda_max)) throw ArgError("svd: the matrix is not finite");
  if (r == 0) return out;
  if (!(out.sel.lambda_max > 0)) {  // a zero matrix: sigma = 0 with unit vectors
    std::vector<double> uh(size_t(m * r), 0.0), vh(size_t(r * n), 0.0);
    for (int64_t j = 0; j < r; ++j) {
      if (j < m) uh[size_t(j * r + j)] = 1.0;
      if (j < n) vh[size_t(j * n + j)] = 1.0;
    }
    TK_HIP(hipMemcpyAsync(U, uh.data(), uh.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    TK_HIP(hipMemcpyAsync(Vt, vh.data(), vh.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
    return out;
  }
  const double* E = bs.u();  // g x g row-major, kept left vectors of X in columns 0..r-1
  const int64_t e_n = bs.u_size();
  // C = E[:, :r]^T X (r x p): Sigma times the other factor
  DBuf<double> Cm(ctx, size_t(r * p)), W(ctx, size_t(p * r)), inv(ctx, size_t(r));
  dense_gemm(ctx, r, p, g, 1.0, E, 1, g, X, p, 1, 0.0, Cm.get(), p, 1);
  std::vector<double> ih(static_cast<size_t>(r));
  for (int64_t j = 0; j < r; ++j) ih[size_t(j)] = out.s[size_t(j)] > 0 ? 1.0 / out.s[size_t(j)] : 0.0;
  inv.upload(ih.data(), ih.size());
  const int grid = int(std::max<int64_t>(1, std::min<int64_t>((r * p + 255) / 256, int64_t(ctx->n_cu) * 8)));
  hipLaunchKernelGGL(scaled_transpose_kernel, dim3(grid), dim3(256), 0, ctx->stream, Cm.get(), inv.get(), r, p,
                     W.get());
  TK_HIP(hipGetLastError());
  // the other factor's columns, orthonormal in descending sigma order (a zero
  // or tiny sigma leaves a column the orthonormalisation reseeds or corrects)
  orthonormalize_columns(ctx, {OrthJob{0, int32_t(p), int32_t(r), 0}}, W.get(), p * r);
  if (rows_side) {
    // U = E[:, :r] (m x r); Vt = W^T (r x n)
    launch_copies(ctx, {CopyJob{0, 0, int32_t(m), int32_t(r), int32_t(g), int32_t(r), 0, 0}}, E, e_n, U, m * r);
    launch_copies(ctx, {CopyJob{0, 0, int32_t(r), int32_t(n), int32_t(r), int32_t(n), 1, 0}}, W.get(), p * r, Vt,
                  r * n);
  } else {
    // Vt = E[:, :r]^T (r x n); U = W (m x r)
    launch_copies(ctx, {CopyJob{0, 0, int32_t(r), int32_t(n), int32_t(g), int32_t(n), 1, 0}}, E, e_n, Vt, r * n);
    launch_copies(ctx, {CopyJob{0, 0, int32_t(m), int32_t(r), int32_t(r), int32_t(r), 0, 0}}, W.get(), p * r, U,
                  m * r);
  }
  ctx->sync();
  return out;
}

}  // namespace tk
