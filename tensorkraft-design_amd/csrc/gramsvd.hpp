// Block-diagonal truncated SVD through Gram matrices, refined once on the
// projected remainder (gramsvd.hip).
//
// For dense blocks X_b (n_b x p_b) the squared singular values and left
// singular vectors come from the eigen-decomposition of G_b = X_b X_b^T.  An
// FP64 Gram matrix resolves lambda = sigma^2 only down to ~1e-15 lambda_max,
// i.e. sigma >= ~3e-8 sigma_max, while the reference keeps every sigma >=
// svd_cutoff * sigma_max with svd_cutoff = 1e-12 (tk-dmrg/src/sweep/mod.rs:151,
// truncation/mod.rs:64-80).  When more singular values than level one resolves
// could be kept, the level-one vectors U1 with lambda >= kTau1 lambda_max are
// projected out, X2 = X - U1 (U1^T X), and the Gram matrix of X2 gives the
// remaining sigma^2 down to ~1e-15 of its own largest value; the kept vectors
// of both levels are then orthonormalised together in descending order.
// Reference call site: truncate_svd -> faer thin_svd
// (tk-dmrg/src/truncation/mod.rs:56-127, tk-linalg/src/device/faer.rs:191-240).
#pragma once
#include <cstdint>
#include <vector>

#include "device.hpp"
#include "tdeig.hpp"

namespace tk {

// X_b: n x p row-major at x_off.
struct GramBlock {
  int64_t x_off;
  int32_t n, p;
};

class GramSvd {
 public:
  static constexpr double kTau1 = 1e-9;  // level-one resolution (relative lambda)
  explicit GramSvd(Context* ctx) : ctx_(ctx), eig1_(ctx), eig2_(ctx) {}
  // Candidate lambda = sigma^2 of every block, descending, at most min(n, p)
  // values each.  Level two runs when lambda_min_rel < kTau1 and fewer than
  // max_keep values were resolved by level one.  lambda_max is the largest
  // value over all blocks.
  std::vector<std::vector<double>> values(const std::vector<GramBlock>& blocks, const double* X, int64_t x_n,
                                          double lambda_min_rel, int64_t max_keep);
  // The first keep[b] left singular vectors of block b as the columns
  // 0..keep[b]-1 of an n_b x n_b row-major matrix at u_offset(b) of u().
  void vectors(const std::vector<int64_t>& keep);
  const double* u() const { return E1_.get(); }
  int64_t u_size() const { return gtot_; }
  int64_t u_offset(size_t b) const { return g_off_[b]; }
  bool refined() const { return refined_; }

 private:
  Context* ctx_;
  SymEigBatch eig1_, eig2_;
  std::vector<GramBlock> blocks_;
  std::vector<int64_t> g_off_, w_off_, k1_, rank_;
  std::vector<int> lvl2_;  // index of the block in the level-two job list, -1 if none
  int64_t gtot_ = 0, wtot_ = 0;
  bool refined_ = false;
  DBuf<double> G1_, E1_, W1_, G2_, E2_, W2_;
};

}  // namespace tk
