// Impurity-model pieces of the DMFT self-consistency loop (dmft.hip).
//
//  * hybridization: Delta(z) of many baths at many complex frequencies in one
//    launch, star form sum_k v_k^2 / (z - eps_k) (BathParameters::
//    hybridization_function, tk-dmft/src/impurity/bath.rs:61-83, with
//    z = omega + i*broadening there) or chain form (the continued fraction of
//    a tight-binding chain, the geometry of the AIM Hamiltonian).
//  * bath_fit: batched Levenberg-Marquardt least-squares fit of star bath
//    parameters to Delta(i omega_n) on the Matsubara axis (the fallback bath
//    update; no reference counterpart).
//  * lanczos_tridiagonalize: the reference discretiser (tk-dmft/src/impurity/
//    discretize.rs:54-274) on a weighted point set; applied to a star bath it
//    gives the equivalent chain exactly.
//  * BroydenMixer: BroydenState::update (tk-dmft/src/loop/mixing.rs:40-150).
//  * krylov_tridiag: Lanczos coefficients of a CSR matrix from a start vector
//    (full reorthogonalisation), the continued fraction of an impurity Green's
//    function.
//  * Mlp: the learned bath proposal, FP32 dense layers (tanh hidden, linear
//    output) with forward, backward and Adam on the device.
#pragma once
#include <cstdint>
#include <vector>

#include "device.hpp"

namespace tk {

enum class BathGeometry : int { Star = 0, Chain = 1 };

// out[(p * n_z + j) * 2 + {0,1}] = Delta_p(z_j); params per problem:
// eps[0..n_bath), v[0..n_bath).
void hybridization(Context* ctx, BathGeometry g, int n_problems, int n_bath, const double* params, int n_z,
                   const double* z_re, const double* z_im, double* out);

// Fits star parameters per problem to delta (n_problems x n_z complex) at
// z_j = i * wn[j]; params holds the start on entry and the fit on exit;
// cost (host, n_problems) = sum_j |Delta_fit - delta|^2.  n_z <= 256,
// n_bath <= 8.
void bath_fit(Context* ctx, int n_problems, int n_bath, int n_z, const double* wn, const double* delta,
              double* params, int max_iter, double* cost);

// Lanczos recursion on the weighted points (x_j, w_j >= 0):
// eps[k] = alpha_k, v[0] = sqrt(sum w), v[k] = beta_k (discretize.rs:114-233).
void lanczos_tridiagonalize(int n, const double* x, const double* w, int n_bath, double* eps, double* v);

class BroydenMixer {
 public:
  explicit BroydenMixer(int history) : max_history_(history) {}
  std::vector<double> update(const std::vector<double>& x, const std::vector<double>& f, double alpha);
  int n_updates() const { return n_updates_; }

 private:
  std::vector<double> prev_x_, prev_r_, jinv_;
  int dim_ = 0, n_updates_ = 0, max_history_;
  bool have_prev_ = false;
};

// alpha (m), beta (m): beta[0] = |v0|, then the Lanczos recurrence
// A q_j = beta_j q_{j-1} + alpha_j q_j + beta_{j+1} q_{j+1}; returns the
// number of steps taken (stops at beta < 1e-13 |A| estimate).
int krylov_tridiag(Context* ctx, int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                   const double* v0, int m, double* alpha, double* beta);

class Mlp {
 public:
  Mlp(Context* ctx, const std::vector<int>& dims, uint64_t seed);
  int n_params() const { return int(n_params_); }
  int in_dim() const { return dims_.front(); }
  int out_dim() const { return dims_.back(); }
  void get_params(float* host) const;
  void set_params(const float* host);
  // y (batch x out) = network(x (batch x in)); device pointers
  void forward(int batch, const float* x, float* y);
  // mean squared error against t and its gradient (kept for adam_step;
  // copied to grad_host when not null)
  float loss_grad(int batch, const float* x, const float* t, float* grad_host);
  void adam_step(float lr, float beta1, float beta2, float eps);

 private:
  void ensure(int batch);
  Context* ctx_;
  std::vector<int> dims_;
  std::vector<int64_t> w_off_, b_off_;
  int64_t n_params_ = 0;
  int batch_cap_ = 0, t_ = 0;
  DBuf<float> p_, g_, m_, v_, act_, dact_;
  std::vector<int64_t> act_off_;
};

}  // namespace tk
