// Device-resident Davidson eigensolver for the lowest eigenpair of H_eff.
//
// Reference: DavidsonSolver::lowest_eigenpair (tk-dmrg/src/eigensolver/mod.rs:325-516):
// projected matrix, lowest Ritz pair, residual r = A y - theta y, diagonal
// preconditioner t_i = r_i / (theta - H_ii) (skipped where |theta - H_ii| <= 1e-14),
// Gram-Schmidt against the basis, raw-residual fallback when the corrected
// vector collapses, restart to the current Ritz vector at max_subspace.
// Differences (stated): the start vector is the caller's (warm start from the
// current two-site tensor), orthogonalisation is classical Gram-Schmidt applied
// twice, and the projected problem is solved by the batched Jacobi kernel on
// the device.
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

}  // namespace tk
