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
// svd_truncated (traits.rs:38-88 with faer.rs:127-240): row-major m x n A;
// U (m x rank) and Vt (rank x n) row-major on the device.
SvdOut svd_truncated(Context* ctx, int64_t m, int64_t n, const double* A, int64_t max_rank, double cutoff,
                     double* U, double* Vt);

// All eigenpairs of a symmetric n x n matrix, ascending (eigh_lowest, traits.rs:134-138).
std::vector<double> eigh(Context* ctx, int64_t n, const double* A, double* evecs);

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
  void spmv(Sym sym, const QIndex& rows, const QIndex& cols, QN flux, const double* a, const double* x,
            double* y) const {
    block_spmv(ctx_, sym, rows, cols, flux, a, x, y);
  }
  Context* context() const { return ctx_; }

 private:
  Context* ctx_;
};

}  // namespace tk
