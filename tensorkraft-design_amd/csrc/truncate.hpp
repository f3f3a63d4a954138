// SVD truncation: the selection rule, the bond-dimension schedule and the
// block-diagonal truncated SVD shared by the two-site split, the dense
// svd_truncated / truncate_svd entry points and the C ABI.
//
// Reference: truncate_svd and BondDimensionSchedule
// (tk-dmrg/src/truncation/mod.rs:56-172), the truncate_svd! macro of faer's
// svd_truncated (tk-linalg/src/device/faer.rs:127-181).
#pragma once
#include <cstdint>
#include <vector>

#include "device.hpp"
#include "gramsvd.hpp"

namespace tk {

// TruncationConfig (truncation/mod.rs:12-30) with the two rules this build
// adds for the sweep (both off at 0):
//   degeneracy_tol: never cut between sigma_k and sigma_k+1 when
//     sigma_k - sigma_k+1 <= degeneracy_tol * sigma_max (the cut moves up to
//     the start of the group, unless that would keep nothing);
//   lambda_floor: sigma^2 < lambda_floor * sigma_max^2 is a numerical zero and
//     is never kept by the cutoff (min_bond can still take it).
struct TruncRule {
  int64_t max_bond = INT64_MAX;
  double cutoff = 0.0;  // relative: keep sigma >= cutoff * sigma_max
  int64_t min_bond = 0;
  double degeneracy_tol = 0.0;
  double lambda_floor = 0.0;
};

struct TruncSelection {
  std::vector<int64_t> keep;  // kept count per block
  int64_t kept = 0, discarded = 0;
  double sum_all = 0.0, sum_kept = 0.0;  // of sigma^2
  double lambda_max = 0.0;
  // TruncationResult::truncation_error (truncation/mod.rs:85-98): relative
  double rel_error() const { return sum_all > 0 ? (sum_all - sum_kept) / sum_all : 0.0; }
  // SvdResult::truncation_error (tk-linalg/src/results.rs:18-23): absolute
  double abs_error() const { return sum_all - sum_kept > 0 ? sum_all - sum_kept : 0.0; }
};

// Global selection over the sigma^2 values of several blocks (each
// descending).  Candidates are ordered by sigma descending, ties by block then
// position; then the reference's loop (truncation/mod.rs:72-82): keep while
// count < max_bond and sigma >= cutoff * sigma_max; then the degeneracy and
// floor rules above; then at least min(min_bond, candidates).
TruncSelection select_truncation(const std::vector<std::vector<double>>& lambda, const TruncRule& rule);

// BondDimensionSchedule (truncation/mod.rs:131-172).
struct BondSchedule {
  std::vector<int64_t> dims;
  // round(d_init * r^i), r = (d_max / d_init)^(1/n), clamped to [1, d_max];
  // n = 0 gives [d_max]
  static BondSchedule warmup(int64_t d_init, int64_t d_max, int64_t n_warmup_sweeps);
  static BondSchedule fixed(int64_t d) { return BondSchedule{{d}}; }
  static BondSchedule custom(std::vector<int64_t> d) { return BondSchedule{std::move(d)}; }
  // the last value for sweeps beyond the schedule (1 when empty)
  int64_t at_sweep(int64_t sweep_index) const;
};

// Truncated SVD of dense blocks X_b (n_b x p_b, row-major at x_off of X)
// through the two-level Gram eigen-decomposition: sigma^2 of every block,
// the selection of `rule`, then the kept left singular vectors of each block
// as columns 0..keep[b)-1 of an n_b x n_b row-major matrix at u_offset(b) of
// u() (the matrix's other columns are scratch).
class BlockSvd {
 public:
  explicit BlockSvd(Context* ctx) : ctx_(ctx), gs_(ctx) {}
  const TruncSelection& run(const std::vector<GramBlock>& blocks, const double* X, int64_t x_n,
                            const TruncRule& rule);
  const std::vector<std::vector<double>>& lambda() const { return lambda_; }
  const double* u() const { return gs_.u(); }
  int64_t u_size() const { return gs_.u_size(); }
  int64_t u_offset(size_t b) const { return gs_.u_offset(b); }
  const TruncSelection& selection() const { return sel_; }

 private:
  Context* ctx_;
  GramSvd gs_;
  std::vector<std::vector<double>> lambda_;
  TruncSelection sel_;
};

struct DenseSvd {
  int64_t rank = 0;
  std::vector<double> s;  // kept singular values, descending
  TruncSelection sel;
};
// Truncated SVD of one row-major m x n matrix (min(m, n) <= 2048): U (m x
// rank) and Vt (rank x n) row-major on the device.  The Gram matrix is taken
// on the smaller side; the other factor is A^T U / sigma (or A V / sigma),
// orthonormalised in descending order.  Every one of the min(m, n) values is
// a candidate (values the Gram levels do not resolve count as zero), so
// min_bond can keep them.
DenseSvd dense_svd_truncate(Context* ctx, int64_t m, int64_t n, const double* A, const TruncRule& rule, double* U,
                            double* Vt);

}  // namespace tk
