// Device-resident Davidson eigensolver for the lowest eigenpair of H_eff.
//
// Reference: DavidsonSolver::lowest_eigenpair (tk-dmrg/src/eigensolver/mod.rs:325-516):
// projected matrix, lowest Ritz pair, residual r = A y - theta y, diagonal
// preconditioner t_i = r_i / (theta - H_ii) (skipped where |theta - H_ii| <= 1e-14),
// Gram-Schmidt against the basis, raw-residual fallback when the corrected
// vector collapses, restart to the current Ritz vector at max_subspace.
// Differences (stated): the start vector is the caller's (warm start from the
// current two-site tensor), orthogonalisation is classical Gram-Schmidt applied
// twice, and the lowest pair of the projected problem comes from a one-
// wavefront kernel (Householder tridiagonalisation, multisection, inverse
// iteration) on the device.
//
// Mixed precision, with a plan precision below FP64, in two places:
// - a cold start (first residual above low_start_rel * |theta|) multiplies new
//   basis vectors with the low-precision kernels until the residual stalls or
//   falls below switch_rel * |theta|, then restarts from the Ritz vector with
//   FP64 matvecs (FP64 residual refinement);
// - once the residual of an FP64 iteration is below low_tail * tol, the
//   remaining corrections are multiplied in low precision without a restart.
//   The projected matrix takes a new vector t's couplings to the earlier
//   columns as (A v_j) . t from the products held (A is symmetric), not as
//   v_j . (A t) from the rounded product, whose error (about eps |A|, eps the
//   low precision's relative error) would exceed the true coupling of t to
//   the start vector; only H_tt uses the rounded product.  The coefficient of
//   t in the Ritz vector is then of the order of the residual over the gap,
//   so the rounding moves the residual by about that coefficient times
//   eps |A|: below tol for low_tail = 1000 in FP32 (eps about 4e-7) up to
//   |A| / gap of about 2500, in split-BF16 (eps about 6e-6) up to about 150.  A residual that stops falling (three iterations
//   without a 0.7 reduction) restarts from the Ritz vector in FP64.
// The residual test is on the Ritz residual of the basis held.
#pragma once
#include <cstdint>

#include "device.hpp"
#include "heff.hpp"

namespace tk {

struct DavidsonOpts {
  int max_subspace = 60;
  int max_iter = 1000;
  double tol = 1e-10;          // absolute 2-norm of the residual
  bool precondition = true;
  double switch_rel = 1e-6;    // low -> FP64 switch when ||r|| < switch_rel * |theta|
  // A low-precision solve starts in FP64 and drops to low precision only
  // when the first residual is above low_start_rel * |theta| (a warm start
  // already that close is refined in FP64 without a basis restart).
  double low_start_rel = 1e-4;
  // low-precision corrections once ||r|| < low_tail * tol (0 or less: off)
  double low_tail = 1000;
};

struct DavidsonResult {
  double eigenvalue = 0;
  double residual = 0;
  int matvecs = 0;
  int matvecs_fp64 = 0;
  int iterations = 0;
  bool converged = false;
};

// x: device vector (length plan.psi().total), the start vector on entry (any
// norm, nonzero), the normalised eigenvector on exit.
DavidsonResult davidson(Context* ctx, HeffPlan& plan, double* x, const DavidsonOpts& opts);

// Generic form over a matvec callback (used by the C ABI for tests):
// matvec(x_dev, y_dev, fp64) computes y = A x.
using MatvecFn = void (*)(const double* x, double* y, int fp64, void* user);
DavidsonResult davidson_generic(Context* ctx, int64_t n, MatvecFn mv, void* user, const double* diag_dev,
                                bool low_precision, double* x, const DavidsonOpts& opts);

// Lanczos with full reorthogonalisation for the lowest eigenpair.
//
// Reference: LanczosSolver::lowest_eigenpair + solve_tridiagonal +
// count_eigenvalues_below (tk-dmrg/src/eigensolver/mod.rs:57-297): three-term
// recurrence alpha_j = v_j.(A v_j), full reorthogonalisation against every
// Krylov vector, beta_j = |w|, stop at beta_j < 1e-15 or at max_krylov, lowest
// eigenvalue of T by Sturm bisection.  Differences (stated): the eigenvector of
// T comes from inverse iteration (the reference leaves it at all ones,
// eigensolver/mod.rs:235-247), the solver stops when the Ritz residual
// estimate |beta_j z_j| is below tol and restarts from the Ritz vector at
// max_krylov (the reference returns after one pass), reorthogonalisation is
// classical Gram-Schmidt applied twice, and the reported residual is the
// explicit |A y - theta y| of the returned pair.  Mixed precision as in the
// Davidson solver: low-precision matvecs until the estimate falls below
// switch_rel * |theta|, then a restart with FP64 matvecs.
struct LanczosOpts {
  int max_krylov = 100;
  int max_restarts = 1000;
  double tol = 1e-10;
  double switch_rel = 1e-6;
};

DavidsonResult lanczos_generic(Context* ctx, int64_t n, MatvecFn mv, void* user, bool low_precision, double* x,
                               const LanczosOpts& opts);
DavidsonResult lanczos(Context* ctx, HeffPlan& plan, double* x, const LanczosOpts& opts);

}  // namespace tk
