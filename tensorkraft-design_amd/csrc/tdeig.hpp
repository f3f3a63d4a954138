// Batched FP64 symmetric eigensolver for the truncated SVD and the Krylov
// subspace problems (tdeig.hip).
//
// Method:
//   1. Householder tridiagonalisation A = Q T Q^T, the rank-2 update of step
//      k fused with the symmetric product of step k+1, so the trailing matrix
//      is read and written once per column; reflectors are kept in A.  One
//      launch per column step over the whole batch while a trailing block is
//      larger than 192, then one workgroup per matrix with the block in
//      registers.
//   2. All eigenvalues of T by Sturm-count multisection (16 lanes per
//      eigenvalue), sorted by construction.
//   3. For the k wanted eigenvalues only: inverse iteration on T (LU with
//      partial pivoting, one thread per vector), blocked classical
//      Gram-Schmidt applied twice in eigenvalue order with Cholesky QR inside
//      each 64-column block (orthonormal to working precision even inside
//      clusters: a block with a dependent column falls back to column-by-column
//      Gram-Schmidt), then back-transformation by Q with the reflectors grouped
//      128 at a time in compact WY form (I - V T V^T), applied as grouped MFMA
//      contractions.
//
// Reference call sites: the Gram eigen-decomposition behind the block-diagonal
// truncated SVD (tk-dmrg/src/truncation/mod.rs:56-127 -> faer thin_svd,
// tk-linalg/src/device/faer.rs:191-240) and the projected matrix of the
// Lanczos/Davidson solvers (symmetric_eigen_small,
// tk-dmrg/src/eigensolver/mod.rs:532-601).
#pragma once
#include <cstdint>
#include <vector>

#include "device.hpp"

namespace tk {

constexpr int kSymEigMaxN = 2048;

// One symmetric n x n matrix, row-major at a_off of A (upper triangle read,
// overwritten by the reflectors); its eigenvalues at w_off of W.
struct SymJob {
  int64_t a_off;
  int64_t w_off;
  int32_t n;
};

// q[i] = a[i] / b[i] (device arrays of n) through the hardware reciprocal
// and one Newton step, the quotient of the Sturm counts and of inverse
// iteration; synchronous.
void fast_quotient(Context* ctx, const double* a, const double* b, double* q, int64_t n);

class SymEigBatch {
 public:
  explicit SymEigBatch(Context* ctx) : ctx_(ctx) {}
  // Eigenvalues of every job into W (descending or ascending).
  void values(const std::vector<SymJob>& jobs, double* A, int64_t a_n, double* W, int64_t w_n, bool descending);
  // The first keep[j] eigenvectors (in W order) of job j as the columns
  // 0..keep[j]-1 of the n x n row-major matrix at a_off of Z (Z != A).
  void vectors(const std::vector<int64_t>& keep, double* Z, int64_t z_n);

 private:
  Context* ctx_;
  std::vector<SymJob> jobs_;
  double* A_ = nullptr;
  double* W_ = nullptr;
  int64_t a_n_ = 0;
  int nmax_ = 0;
  DBuf<double> de_;   // per job: d (n), e (n), tau (n) at 3 * w_off
  DBuf<double> tn_;   // per job: Gershgorin bound of |T|
};

// Orthonormalise the k columns of each n x k row-major matrix (leading
// dimension k) at off of Z, in column order (classical Gram-Schmidt applied
// twice, blocked), in place.  Columns before `first` are orthonormal already
// and stay as they are.
struct OrthJob {
  int64_t off;
  int32_t n, k;
  int32_t first = 0;
};
void orthonormalize_columns(Context* ctx, const std::vector<OrthJob>& jobs, double* Z, int64_t z_n);

// Lowest eigenpair of the m x m symmetric matrix whose lower triangle is
// stored row-wise in H (H[i*ld + j], j <= i), m <= 64: one wavefront,
// everything in LDS.  c (m) and *theta are device outputs.
void lowest_eigpair(Context* ctx, const double* H, int ld, int m, double* c, double* theta);

// Lowest eigenpair of the symmetric tridiagonal with diagonal alpha (m) and
// off-diagonal beta (m-1), m <= 512, all on the device: z (m, unit norm),
// *theta, and when est is not null *est = |*beta_next * z[m-1]| (the Lanczos
// residual estimate of the Ritz pair).  One wavefront.
void tridiag_lowest(Context* ctx, const double* alpha, const double* beta, int m, const double* beta_next,
                    double* z, double* theta, double* est);

}  // namespace tk
