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
// Mixed precision: while the plan's precision is below FP64, new basis vectors
// are multiplied with the low-precision kernels; once the residual stalls or
// falls below `switch_rel * |theta|`, the solver restarts from the Ritz vector
// with FP64 matvecs (FP64 residual refinement) and finishes with the FP64
// residual test, so the returned eigenvalue is an FP64 Rayleigh quotient.
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
