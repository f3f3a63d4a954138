// Real-valued Krylov exponential exp(alpha A) v on the device.
//
// Reference: exp_krylov_f64 + mat_exp_times_vec (tk-dmrg/src/tdvp/mod.rs:197-382):
// m = min(krylov_dim, n) Arnoldi steps from v / |v| with a happy-breakdown
// exit at h_{j+1,j} < tol, then exp(alpha H_m) (|v| e_1) by a Taylor series
// (vector form when the scaling exponent s is 0, else the matrix form on
// alpha H_m 2^-s, squared s times) and the result V_m c.  A vector with
// |v| < 1e-30 maps to zero.  Differences (stated): the Arnoldi column is
// orthogonalised by classical Gram-Schmidt applied twice (the reference uses
// one modified pass; both give the same Hessenberg matrix up to rounding) and
// the small exponential runs in one workgroup on the device.
#pragma once
#include <cstdint>

#include "davidson.hpp"
#include "device.hpp"
#include "heff.hpp"

namespace tk {

// Largest Krylov dimension (three m x m FP64 matrices in LDS).
constexpr int kExpKrylovMax = 48;

struct ExpKrylovResult {
  int krylov_dim = 0;   // Arnoldi vectors used (0 for the zero vector)
  int matvecs = 0;
  double norm_in = 0;   // |v|
  double norm_out = 0;  // |exp(alpha A) v| = |c| (orthonormal basis)
};

// out = exp(alpha A) v; v and out are device vectors of length n (may alias).
// adaptive_tol > 0 adds an early stop the reference does not have: after
// Arnoldi step m the small exponential c = exp(alpha H_m) |v| e_1 is formed
// and the iteration ends once h_{m+1,m} |c_m| <= adaptive_tol |c| (the size
// of the next Krylov direction's contribution); krylov_dim is then a maximum.
ExpKrylovResult exp_krylov_generic(Context* ctx, int64_t n, MatvecFn mv, void* user, const double* v, double alpha,
                                   int krylov_dim, double tol, double* out, double adaptive_tol = 0.0);
// Over an H_eff plan, with FP64 matvecs.
ExpKrylovResult exp_krylov(Context* ctx, HeffPlan& plan, const double* v, double alpha, int krylov_dim, double tol,
                           double* out, double adaptive_tol = 0.0);

}  // namespace tk
