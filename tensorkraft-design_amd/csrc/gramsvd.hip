// Two-level Gram SVD of dense blocks; the method is described in gramsvd.hpp.
#include "gramsvd.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "contract.hpp"
#include "kernels.hpp"

namespace tk {
namespace {

void run(Context* ctx, ContractList& cl, int64_t a_n, int64_t b_n, int64_t c_n, const double* A, const double* B,
         double* C) {
  if (cl.outs.empty()) return;
  cl.finalize();
  cl.validate(a_n, b_n, c_n);
  DeviceContractList d;
  d.upload(ctx, cl);
  launch_contract(ctx, d, A, B, C);
}

// G_b = X_b X_b^T for the selected blocks
void gram(Context* ctx, const std::vector<GramBlock>& blocks, const std::vector<int64_t>& g_off,
          const std::vector<char>& sel, const double* X, int64_t x_n, double* G, int64_t g_n) {
  ContractList cl;
  cl.flags = kTransB;
  for (size_t b = 0; b < blocks.size(); ++b) {
    if (!sel[b]) continue;
    const GramBlock& k = blocks[b];
    cl.begin_out(g_off[b], k.n, k.n, k.n);
    cl.begin_term(k.x_off, k.p, k.p);
    cl.add_src(k.x_off, k.p, 1.0);
  }
  run(ctx, cl, x_n, x_n, g_n, X, X, G);
}

}  // namespace

std::vector<std::vector<double>> GramSvd::values(const std::vector<GramBlock>& blocks, const double* X, int64_t x_n,
                                                 double lambda_min_rel, int64_t max_keep) {
  blocks_ = blocks;
  const size_t nb = blocks.size();
  g_off_.assign(nb, 0);
  w_off_.assign(nb, 0);
  k1_.assign(nb, 0);
  rank_.assign(nb, 0);
  lvl2_.assign(nb, -1);
  gtot_ = 0;
  wtot_ = 0;
  refined_ = false;
  for (size_t b = 0; b < nb; ++b) {
    const GramBlock& k = blocks[b];
    if (k.n < 1 || k.p < 1 || k.n > kSymEigMaxN) throw ArgError("gram svd: block shape outside 1..2048 rows");
    if (k.x_off < 0 || k.x_off + int64_t(k.n) * k.p > x_n) throw std::out_of_range("gram svd: block outside X");
    g_off_[b] = gtot_;
    w_off_[b] = wtot_;
    gtot_ += int64_t(k.n) * k.n;
    wtot_ += k.n;
    rank_[b] = std::min(k.n, k.p);
  }
  std::vector<std::vector<double>> out(nb);
  if (nb == 0) return out;
  G1_.reset(ctx_, size_t(gtot_));
  E1_.reset(ctx_, size_t(gtot_));
  W1_.reset(ctx_, size_t(wtot_));
  gram(ctx_, blocks, g_off_, std::vector<char>(nb, 1), X, x_n, G1_.get(), gtot_);
  std::vector<SymJob> jobs1;
  for (size_t b = 0; b < nb; ++b) jobs1.push_back({g_off_[b], w_off_[b], blocks[b].n});
  eig1_.values(jobs1, G1_.get(), gtot_, W1_.get(), wtot_, true);
  std::vector<double> lam1(static_cast<size_t>(wtot_));
  W1_.download(lam1.data(), lam1.size());
  ctx_->sync();
  double lmax = 0;
  for (size_t b = 0; b < nb; ++b) lmax = std::max(lmax, lam1[w_off_[b]]);
  int64_t k1_total = 0;
  bool remainder = false;
  for (size_t b = 0; b < nb; ++b) {
    for (int64_t j = 0; j < rank_[b]; ++j) {
      const double l = lam1[w_off_[b] + j];
      out[b].push_back(l);
      if (l >= kTau1 * lmax) ++k1_[b];
    }
    k1_total += k1_[b];
    remainder = remainder || rank_[b] > k1_[b];
  }
  if (!(lmax > 0) || lambda_min_rel >= kTau1 || k1_total >= max_keep || !remainder) return out;
  // level two: X2 = X - U1 (U1^T X), Gram of X2
  refined_ = true;
  const int ph = ctx_->timer.begin(ctx_->stream);
  eig1_.vectors(k1_, E1_.get(), gtot_);
  std::vector<char> sel(nb, 0);
  std::vector<int64_t> c_off(nb, 0);
  int64_t ctot = 0;
  for (size_t b = 0; b < nb; ++b)
    if (rank_[b] > k1_[b]) {
      sel[b] = 1;
      c_off[b] = ctot;
      ctot += k1_[b] * blocks[b].p;
    }
  DBuf<double> X2(ctx_, size_t(x_n));
  TK_HIP(hipMemcpyAsync(X2.get(), X, size_t(x_n) * sizeof(double), hipMemcpyDeviceToDevice, ctx_->stream));
  if (ctot > 0) {
    DBuf<double> C1(ctx_, size_t(ctot));
    ContractList c1;  // C1 = U1^T X (k1 x p)
    c1.flags = kTransA;
    ContractList up;  // X2 = X2 - U1 C1
    up.beta = 1.0;
    for (size_t b = 0; b < nb; ++b) {
      if (!sel[b] || k1_[b] == 0) continue;
      const GramBlock& k = blocks[b];
      c1.begin_out(c_off[b], int(k1_[b]), k.p, k.p);
      c1.begin_term(k.x_off, k.p, k.n);
      c1.add_src(g_off_[b], k.n, 1.0);
      up.begin_out(k.x_off, k.n, k.p, k.p);
      up.begin_term(c_off[b], k.p, int(k1_[b]));
      up.add_src(g_off_[b], k.n, -1.0);
    }
    run(ctx_, c1, gtot_, x_n, ctot, E1_.get(), X, C1.get());
    run(ctx_, up, gtot_, ctot, x_n, E1_.get(), C1.get(), X2.get());
  }
  G2_.reset(ctx_, size_t(gtot_));
  W2_.reset(ctx_, size_t(wtot_));
  gram(ctx_, blocks, g_off_, sel, X2.get(), x_n, G2_.get(), gtot_);
  std::vector<SymJob> jobs2;
  for (size_t b = 0; b < nb; ++b)
    if (sel[b]) {
      lvl2_[b] = int(jobs2.size());
      jobs2.push_back({g_off_[b], w_off_[b], blocks[b].n});
    }
  eig2_.values(jobs2, G2_.get(), gtot_, W2_.get(), wtot_, true);
  std::vector<double> lam2(static_cast<size_t>(wtot_));
  W2_.download(lam2.data(), lam2.size());
  ctx_->sync();
  for (size_t b = 0; b < nb; ++b) {
    if (!sel[b]) continue;
    out[b].resize(size_t(k1_[b]));
    const double top = lam2[w_off_[b]];
    for (int64_t j = 0; j < rank_[b] - k1_[b]; ++j) {
      const double l = lam2[w_off_[b] + j];
      if (!(top > 0) || l < 1e-15 * top || !(l > 0)) break;
      out[b].push_back(l);
    }
  }
  ctx_->timer.end(ph, ctx_->stream, "phase_svd_level2", 0.0, 0.0, 1);
  return out;
}

void GramSvd::vectors(const std::vector<int64_t>& keep) {
  const size_t nb = blocks_.size();
  if (keep.size() != nb) throw ArgError("gram svd: one kept count per block");
  for (size_t b = 0; b < nb; ++b)
    if (keep[b] < 0 || keep[b] > rank_[b]) throw ArgError("gram svd: kept count outside 0..rank");
  if (!refined_) {
    eig1_.vectors(keep, E1_.get(), gtot_);
    return;
  }
  // level-one vectors (k1 columns) are already in E1
  std::vector<int64_t> keep2;
  std::vector<int64_t> z_off(nb, 0);
  int64_t ztot = 0;
  bool any = false;
  for (size_t b = 0; b < nb; ++b) {
    const int64_t k2 = std::max<int64_t>(0, keep[b] - k1_[b]);
    if (lvl2_[b] >= 0) keep2.push_back(k2);
    else if (k2 > 0) throw ArgError("gram svd: kept count beyond the resolved values");
    if (k2 > 0) {
      z_off[b] = ztot;
      ztot += int64_t(blocks_[b].n) * keep[b];
      any = true;
    }
  }
  if (!any) return;
  E2_.reset(ctx_, size_t(gtot_));
  eig2_.vectors(keep2, E2_.get(), gtot_);
  DBuf<double> Z(ctx_, size_t(ztot));
  std::vector<CopyJob> from1, from2, back;
  std::vector<OrthJob> oj;
  for (size_t b = 0; b < nb; ++b) {
    const int64_t k2 = std::max<int64_t>(0, keep[b] - k1_[b]);
    if (k2 == 0) continue;
    const int32_t n = blocks_[b].n;
    const int32_t kk = int32_t(keep[b]);
    if (k1_[b] > 0) from1.push_back({g_off_[b], z_off[b], n, int32_t(k1_[b]), n, kk, 0, 0});
    from2.push_back({g_off_[b], z_off[b] + k1_[b], n, int32_t(k2), n, kk, 0, 0});
    back.push_back({z_off[b], g_off_[b], n, kk, kk, n, 0, 0});
    oj.push_back({z_off[b], n, kk, int32_t(k1_[b])});  // U1 columns are orthonormal already
  }
  launch_copies(ctx_, from1, E1_.get(), gtot_, Z.get(), ztot);
  launch_copies(ctx_, from2, E2_.get(), gtot_, Z.get(), ztot);
  orthonormalize_columns(ctx_, oj, Z.get(), ztot);
  launch_copies(ctx_, back, Z.get(), ztot, E1_.get(), gtot_);
}

}  // namespace tk
