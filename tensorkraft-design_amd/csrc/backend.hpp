// Host-side counterparts of the reference backend traits.
//
//   trait LinAlgBackend<T>        tk-linalg/src/traits.rs:25-184
//   trait SparseLinAlgBackend<T,Q> tk-linalg/src/traits.rs:190-205
//   DeviceAPI<D, S>               tk-linalg/src/device/mod.rs:30-99
//
// The free functions are the device implementations; HipBackend bundles them
// behind the trait-shaped interface that C++ callers (and, through
// include/tk_hip.h, a Rust DeviceHip) program against.  Everything is FP64 and
// device-resident; shape errors throw ArgError (the C ABI maps them to
// TK_ERR_INVALID_ARG, the DimensionMismatch / NonContiguousForBlas cases of
// LinAlgError, tk-linalg/src/error.rs:4-39).
#pragma once
#include <cstdint>
#include <vector>

#include "contract.hpp"
#include "device.hpp"
#include "sym.hpp"

namespace tk {

// C = alpha op(A) op(B) + beta C with MatRef-style strides (traits.rs:114-121).
void dense_gemm(Context* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t rsA,
                int64_t csA, const double* B, int64_t rsB, int64_t csB, double beta, double* C, int64_t rsC,
                int64_t csC);

// dense_gemm with its work list in a static device array instead of an
// allocation (noalloc.hip); lists above that array's size are an ArgError.
void dense_gemm_noalloc(Context* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t rsA,
                        int64_t csA, const double* B, int64_t rsB, int64_t csB, double beta, double* C, int64_t rsC,
                        int64_t csC);

// The work list of one dense_gemm call (finalized and validated against the
// operand extents); false when m or n is 0 (nothing to do).
bool dense_gemm_list(int64_t m, int64_t n, int64_t k, double alpha, int64_t rsA, int64_t csA, int64_t rsB,
                     int64_t csB, double beta, int64_t rsC, int64_t csC, ContractList* cl);

// Rank-2 block_gemm (faer.rs:390-575): output holds the blocks that receive at
// least one sector task, keys ascending; equal output keys accumulate in the
// fixed order (flops descending, key_a, key_b).
class BlockGemmPlan {
 public:
  BlockGemmPlan(Sym sym, const QIndex& a_rows, const QIndex& a_cols, QN a_flux, const QIndex& b_rows,
                const QIndex& b_cols, QN b_flux);
  const std::vector<uint64_t>& keys() const { return keys_; }
  const std::vector<int64_t>& offsets() const { return offsets_; }
  int64_t total() const { return total_; }
  double flops() const { return flops_; }
  void run(Context* ctx, const double* a, const double* b, double* c) const;

 private:
  Sym sym_;
  BlockTable a_, b_;
  std::vector<uint64_t> keys_;
  std::vector<int64_t> offsets_;
  int64_t total_ = 0;
  double flops_ = 0;
  ContractList list_;
};

// spmv (faer.rs:336-388): y = A x over the full row/column index ranges; y is
// overwritten (rows without blocks become zero).
void block_spmv(Context* ctx, Sym sym, const QIndex& rows, const QIndex& cols, QN flux, const double* a,
                const double* x, double* y);

struct SvdOut {
  int64_t rank = 0;
  std::vector<double> s;
  double truncation_error = 0;  // sum of discarded sigma^2 (results.rs:18-23)
};
// svd_truncated (traits.rs:38-88 with faer.rs:127-240): row-major m x n A,
// min(m, n) <= 2048; U (m x rank) and Vt (rank x n) row-major on the device.
// Every sigma >= cutoff * sigma_max is resolved (two Gram levels).
SvdOut svd_truncated(Context* ctx, int64_t m, int64_t n, const double* A, int64_t max_rank, double cutoff,
                     double* U, double* Vt);

// truncate_svd (tk-dmrg/src/truncation/mod.rs:56-127): the full SVD with the
// TruncationConfig rule (max_bond_dim, relative svd_cutoff, min_bond_dim);
// truncation_error relative (1 - kept / total sigma^2).
struct TruncateOut {
  int64_t bond_dim_new = 0, n_discarded = 0;
  std::vector<double> s;
  double truncation_error = 0;
};
TruncateOut truncate_svd(Context* ctx, int64_t rows, int64_t cols, const double* A, int64_t max_bond_dim,
                         double svd_cutoff, int64_t min_bond_dim, double* U, double* Vt);

// qr (traits.rs:148; faer.rs:292-324): thin Householder QR of a row-major m x n
// A: Q (m x k) with orthonormal columns and upper-triangular R (k x n),
// k = min(m, n), both row-major on the device; A is not modified.
void qr(Context* ctx, int64_t m, int64_t n, const double* A, double* Q, double* R);

// regularized_svd_inverse (traits.rs:150-183): V diag(s / (s^2 + delta^2)) U^T
// for U (m x k), Vt (k x n) on the device, s on the host; out is n x m.
void regularized_svd_inverse(Context* ctx, int64_t m, int64_t n, int64_t k, const double* s, const double* U,
                             const double* Vt, double delta, double* out);

// All eigenpairs of a symmetric n x n matrix, ascending (eigh_lowest, traits.rs:134-138).
std::vector<double> eigh(Context* ctx, int64_t n, const double* A, double* evecs);
// eigh_lowest (traits.rs:134-138, faer.rs:243-290): the k lowest eigenpairs,
// eigenvalues ascending, eigenvectors as the columns of a row-major n x k;
// k > n is EighKTooLarge (an ArgError here).
std::vector<double> eigh_lowest(Context* ctx, int64_t n, const double* A, int64_t k, double* evecs);

// The trait surface as one object (DeviceAPI<HipBackend, HipBackend>).
class HipBackend {
 public:
  explicit HipBackend(Context* ctx) : ctx_(ctx) {}
  void gemm(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t rsA, int64_t csA,
            const double* B, int64_t rsB, int64_t csB, double beta, double* C, int64_t rsC, int64_t csC) const {
    dense_gemm(ctx_, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC);
  }
  SvdOut svd_truncated(int64_t m, int64_t n, const double* A, int64_t max_rank, double cutoff, double* U,
                       double* Vt) const {
    return tk::svd_truncated(ctx_, m, n, A, max_rank, cutoff, U, Vt);
  }
  std::vector<double> eigh(int64_t n, const double* A, double* evecs) const { return tk::eigh(ctx_, n, A, evecs); }
  std::vector<double> eigh_lowest(int64_t n, const double* A, int64_t k, double* evecs) const {
    return tk::eigh_lowest(ctx_, n, A, k, evecs);
  }
  void qr(int64_t m, int64_t n, const double* A, double* Q, double* R) const { tk::qr(ctx_, m, n, A, Q, R); }
  void spmv(Sym sym, const QIndex& rows, const QIndex& cols, QN flux, const double* a, const double* x,
            double* y) const {
    block_spmv(ctx_, sym, rows, cols, flux, a, x, y);
  }
  Context* context() const { return ctx_; }

 private:
  Context* ctx_;
};

}  // namespace tk
