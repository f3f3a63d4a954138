// Impurity-model kernels and host pieces of the DMFT loop; see dmft.hpp.
#include "dmft.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>

#include "contract.hpp"
#include "kernels.hpp"

namespace tk {
namespace {

// ---------------------------------------------------------------------------
// Delta(z): star sum_k v_k^2/(z - eps_k), or the chain continued fraction
// v_0^2/(z - eps_0 - v_1^2/(z - eps_1 - ...)).  Complex division of a real
// numerator c by (a + ib) as (c a, -c b) / (a^2 + b^2).
__global__ __launch_bounds__(256) void hybridization_kernel(int geom, int np, int nb,
                                                            const double* __restrict__ params, int nz,
                                                            const double* __restrict__ zr,
                                                            const double* __restrict__ zi,
                                                            double* __restrict__ out) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= int64_t(np) * nz) return;
  const int p = int(g / nz), j = int(g - int64_t(p) * nz);
  const double* eps = params + int64_t(p) * 2 * nb;
  const double* v = eps + nb;
  const double x = zr[j], y = zi[j];
  double re = 0, im = 0;
  if (geom == 0) {
    for (int k = 0; k < nb; ++k) {
      const double a = x - eps[k], b = y, c = v[k] * v[k];
      const double den = a * a + b * b;
      re += (c * a) / den;
      im += (-c * b) / den;
    }
  } else {
    for (int k = nb - 1; k >= 0; --k) {
      const double a = x - eps[k] - re, b = y - im, c = v[k] * v[k];
      const double den = a * a + b * b;
      re = (c * a) / den;
      im = (-c * b) / den;
    }
  }
  out[2 * g] = re;
  out[2 * g + 1] = im;
}

// ---------------------------------------------------------------------------
// Batched Levenberg-Marquardt fit of star parameters theta = (eps, v) to
// delta(i wn_j): one workgroup per problem, one thread per frequency.  The
// residual rows (real and imaginary part per frequency) and the Jacobian are
// staged in LDS; J^T J and J^T r are summed in a fixed order, the damped
// normal equations (Jacobi-scaled, Nielsen's damping update) solved by
// Cholesky on one thread.
constexpr int kFitNT = 256;
constexpr int kFitMaxP = 16;

__device__ void fit_rows(int nb, const double* th, double w, double* re, double* im, double* jre, double* jim) {
  double sr = 0, si = 0;
  for (int k = 0; k < nb; ++k) {
    // q = 1 / (i w - eps) = (a - i b) / (a^2 + b^2), a = -eps, b = w
    const double a = -th[k], b = w, den = a * a + b * b;
    const double qr = a / den, qi = -b / den;
    const double v = th[nb + k], v2 = v * v;
    sr += v2 * qr;
    si += v2 * qi;
    if (jre) {
      // d/d eps = v^2 q^2, d/d v = 2 v q
      const double q2r = qr * qr - qi * qi, q2i = 2 * qr * qi;
      jre[k] = v2 * q2r;
      jim[k] = v2 * q2i;
      jre[nb + k] = 2 * v * qr;
      jim[nb + k] = 2 * v * qi;
    }
  }
  *re = sr;
  *im = si;
}

__global__ __launch_bounds__(kFitNT) void bath_fit_kernel(int nb, int nz, const double* __restrict__ wn,
                                                          const double* __restrict__ delta,
                                                          double* __restrict__ params, int max_iter,
                                                          double* __restrict__ cost_out) {
  const int p = blockIdx.x, tid = threadIdx.x, np = 2 * nb, rows = 2 * nz;
  extern __shared__ double sm[];
  double* J = sm;                         // rows x np
  double* r = J + rows * np;              // rows
  double* part = r + rows;                // kFitNT
  __shared__ double th[kFitMaxP], trial[kFitMaxP], A[kFitMaxP * kFitMaxP], g[kFitMaxP];
  __shared__ double cost_s, lam_s, nu_s, pred_s;
  __shared__ int stop_s;
  const double* dt = delta + int64_t(p) * nz * 2;
  double* par = params + int64_t(p) * np;
  if (tid < np) th[tid] = par[tid];
  if (tid == 0) {
    lam_s = -1.0;  // set from diag(J^T J) at the first iteration
    nu_s = 2.0;
    stop_s = 0;
  }
  __syncthreads();
  double jre[kFitMaxP], jim[kFitMaxP];
  for (int it = 0; it < max_iter; ++it) {
    // residuals and Jacobian at th
    if (tid < nz) {
      double re, im;
      fit_rows(nb, th, wn[tid], &re, &im, jre, jim);
      r[2 * tid] = re - dt[2 * tid];
      r[2 * tid + 1] = im - dt[2 * tid + 1];
      for (int c = 0; c < np; ++c) {
        J[(2 * tid) * np + c] = jre[c];
        J[(2 * tid + 1) * np + c] = jim[c];
      }
    }
    __syncthreads();
    const int nu = np * (np + 1) / 2;
    if (tid < nu) {
      int a = 0, rem = tid;
      while (rem >= np - a) {
        rem -= np - a;
        ++a;
      }
      const int b = a + rem;
      double s = 0;
      for (int q = 0; q < rows; ++q) s += J[q * np + a] * J[q * np + b];
      A[a * kFitMaxP + b] = s;
      A[b * kFitMaxP + a] = s;
    } else if (tid < nu + np) {
      const int a = tid - nu;
      double s = 0;
      for (int q = 0; q < rows; ++q) s += J[q * np + a] * r[q];
      g[a] = s;
    } else if (tid == nu + np) {
      double s = 0;
      for (int q = 0; q < rows; ++q) s += r[q] * r[q];
      cost_s = s;
    }
    __syncthreads();
    if (tid == 0) {
      // Jacobi-scaled normal equations (S A S + lam I) y = -S g, d = S y with
      // S = diag(A)^(-1/2), by Cholesky; the trial point into `trial`
      double M[kFitMaxP * kFitMaxP], sc[kFitMaxP];
      double dmax = 0;
      for (int a = 0; a < np; ++a) {
        const double daa = A[a * kFitMaxP + a];
        sc[a] = daa > 0 ? 1.0 / sqrt(daa) : 1.0;
        dmax = fmax(dmax, daa);
      }
      if (lam_s < 0) lam_s = 1e-3 * fmax(dmax, 1e-300);
      for (int a = 0; a < np; ++a)
        for (int b = 0; b < np; ++b)
          M[a * kFitMaxP + b] = sc[a] * A[a * kFitMaxP + b] * sc[b] + (a == b ? lam_s + 1e-15 : 0.0);
      bool ok = true;
      for (int a = 0; a < np && ok; ++a) {
        double s = M[a * kFitMaxP + a];
        for (int c = 0; c < a; ++c) s -= M[a * kFitMaxP + c] * M[a * kFitMaxP + c];
        if (!(s > 0)) {
          ok = false;
          break;
        }
        const double l = sqrt(s);
        M[a * kFitMaxP + a] = l;
        for (int b = a + 1; b < np; ++b) {
          double t = M[b * kFitMaxP + a];
          for (int c = 0; c < a; ++c) t -= M[b * kFitMaxP + c] * M[a * kFitMaxP + c];
          M[b * kFitMaxP + a] = t / l;
        }
      }
      double y[kFitMaxP];
      for (int a = 0; a < np; ++a) {
        double s = ok ? -sc[a] * g[a] : 0.0;
        if (ok)
          for (int c = 0; c < a; ++c) s -= M[a * kFitMaxP + c] * y[c];
        y[a] = ok ? s / M[a * kFitMaxP + a] : 0.0;
      }
      for (int a = np - 1; a >= 0; --a) {
        double s = y[a];
        if (ok)
          for (int c = a + 1; c < np; ++c) s -= M[c * kFitMaxP + a] * y[c];
        y[a] = ok ? s / M[a * kFitMaxP + a] : 0.0;
      }
      // predicted decrease c - |r + J d|^2 = -2 d.g - d.A d
      double pred = 0;
      for (int a = 0; a < np; ++a) {
        const double da = sc[a] * y[a];
        trial[a] = th[a] + da;
        double ad = 0;
        for (int b = 0; b < np; ++b) ad += A[a * kFitMaxP + b] * sc[b] * y[b];
        pred -= 2.0 * da * g[a] + da * ad;
      }
      pred_s = ok ? pred : 0.0;
    }
    __syncthreads();
    // cost at the trial point
    double cp = 0;
    if (tid < nz) {
      double re, im;
      fit_rows(nb, trial, wn[tid], &re, &im, nullptr, nullptr);
      const double a = re - dt[2 * tid], b = im - dt[2 * tid + 1];
      cp = a * a + b * b;
    }
    part[tid] = cp;
    __syncthreads();
    if (tid == 0) {
      double s = 0;
      for (int q = 0; q < kFitNT; ++q) s += part[q];
      const double c0 = cost_s;
      // Nielsen's damping update on the gain ratio rho = actual / predicted
      const double rho = pred_s > 0 ? (c0 - s) / pred_s : -1.0;
      if (rho > 0 && s < c0) {
        for (int a = 0; a < np; ++a) th[a] = trial[a];
        const double q = 2.0 * rho - 1.0;
        lam_s *= fmax(1.0 / 3.0, 1.0 - q * q * q);
        nu_s = 2.0;
        if (c0 - s <= 1e-15 * c0 || s < 1e-30) stop_s = 1;
        cost_s = s;
      } else {
        lam_s *= nu_s;
        nu_s *= 2.0;
        if (!(lam_s < 1e30)) stop_s = 1;
      }
    }
    __syncthreads();
    if (stop_s) break;
  }
  // final cost at th
  double cp = 0;
  if (tid < nz) {
    double re, im;
    fit_rows(nb, th, wn[tid], &re, &im, nullptr, nullptr);
    const double a = re - dt[2 * tid], b = im - dt[2 * tid + 1];
    cp = a * a + b * b;
  }
  part[tid] = cp;
  __syncthreads();
  if (tid == 0) {
    double s = 0;
    for (int q = 0; q < kFitNT; ++q) s += part[q];
    cost_out[p] = s;
  }
  if (tid < np) par[tid] = th[tid];
}

// ---------------------------------------------------------------------------
// y = A x for a CSR matrix, one thread per row.
__global__ __launch_bounds__(256) void csr_spmv_kernel(int64_t n, const int64_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const double* __restrict__ val, const double* __restrict__ x,
                                                       double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) s += val[q] * x[col[q]];
  y[i] = s;
}

// ---------------------------------------------------------------------------
// MLP elementwise pieces (FP32).
// z (rows x cols) += b (per column); tanh when act
__global__ __launch_bounds__(256) void bias_act_kernel(int64_t rows, int cols, const float* __restrict__ b,
                                                       float* __restrict__ z, int act) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= rows * cols) return;
  const float v = z[g] + b[g % cols];
  z[g] = act ? tanhf(v) : v;
}
// g = 2 (y - t) / (rows * cols); per-block partial sums of (y - t)^2
__global__ __launch_bounds__(256) void mse_grad_kernel(int64_t n, float scale, const float* __restrict__ y,
                                                       const float* __restrict__ t, float* __restrict__ g,
                                                       float* __restrict__ part) {
  __shared__ float red[256];
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float d = 0;
  if (i < n) {
    d = y[i] - t[i];
    g[i] = 2.0f * scale * d;
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// db[c] = sum_r g[r][c] in row order
__global__ __launch_bounds__(256) void colsum_kernel(int rows, int cols, const float* __restrict__ g,
                                                     float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0;
  for (int r = 0; r < rows; ++r) s += g[int64_t(r) * cols + c];
  db[c] = s;
}
// g *= 1 - a^2 (tanh derivative)
__global__ __launch_bounds__(256) void tanh_back_kernel(int64_t n, const float* __restrict__ a, float* __restrict__ g) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) g[i] *= 1.0f - a[i] * a[i];
}
__global__ __launch_bounds__(256) void adam_kernel(int64_t n, float lr, float b1, float b2, float eps, float c1,
                                                   float c2, const float* __restrict__ g, float* __restrict__ m,
                                                   float* __restrict__ v, float* __restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks_for(int64_t n, int nt = 256) { return unsigned((n + nt - 1) / nt); }

// FP32 C (m x n, ldc) = op(A) op(B) through the grouped contraction kernel.
void gemm_f32(Context* ctx, int flags, int m, int n, int k, const float* A, int64_t a_off, int lda, int64_t a_n,
              const float* B, int64_t b_off, int ldb, int64_t b_n, float* C, int64_t c_off, int ldc, int64_t c_n) {
  ContractList cl;
  cl.flags = flags;
  cl.begin_out(c_off, m, n, ldc);
  cl.begin_term(b_off, ldb, k);
  cl.add_src(a_off, lda, 1.0);
  cl.finalize();
  cl.validate(a_n, b_n, c_n);
  DeviceContractList d;
  d.upload(ctx, cl);
  launch_contract(ctx, d, A, B, C);
  ctx->sync();  // the list's tables are released on return
}

}  // namespace

void hybridization(Context* ctx, BathGeometry g, int n_problems, int n_bath, const double* params, int n_z,
                   const double* z_re, const double* z_im, double* out) {
  if (n_problems < 0 || n_bath < 0 || n_z < 0) throw ArgError("hybridization: negative size");
  const int64_t n = int64_t(n_problems) * n_z;
  if (n == 0) return;
  hipLaunchKernelGGL(hybridization_kernel, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, int(g), n_problems,
                     n_bath, params, n_z, z_re, z_im, out);
  TK_HIP(hipGetLastError());
  ctx->sync();
}

void bath_fit(Context* ctx, int n_problems, int n_bath, int n_z, const double* wn, const double* delta,
              double* params, int max_iter, double* cost) {
  if (n_bath < 1 || 2 * n_bath > kFitMaxP) throw ArgError("bath fit: n_bath outside 1..8");
  if (n_z < 1 || n_z > kFitNT) throw ArgError("bath fit: n_z outside 1..256");
  if (n_problems < 1) return;
  const size_t lds = (size_t(2 * n_z) * (2 * n_bath) + 2 * n_z + kFitNT) * sizeof(double);
  TK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&bath_fit_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  DBuf<double> c(ctx, size_t(n_problems));
  hipLaunchKernelGGL(bath_fit_kernel, dim3(n_problems), dim3(kFitNT), lds, ctx->stream, n_bath, n_z, wn, delta,
                     params, max_iter, c.get());
  TK_HIP(hipGetLastError());
  c.download(cost, size_t(n_problems));
  ctx->sync();
}

void lanczos_tridiagonalize(int n, const double* x, const double* w, int n_bath, double* eps, double* v) {
  for (int k = 0; k < n_bath; ++k) eps[k] = v[k] = 0.0;
  if (n_bath == 0 || n == 0) return;
  double wt = 0;
  for (int j = 0; j < n; ++j) wt += w[j];
  if (wt < 2.220446049250313e-16) return;
  const double v1 = std::sqrt(wt);
  std::vector<double> prev(static_cast<size_t>(n), 0.0), cur(static_cast<size_t>(n), 1.0 / v1),
      nxt(static_cast<size_t>(n));
  const int steps = std::min(n_bath, n);
  std::vector<double> al, be{v1};
  for (int k = 0; k < steps; ++k) {
    double a = 0;
    for (int j = 0; j < n; ++j) a += w[j] * x[j] * cur[j] * cur[j];
    al.push_back(a);
    if (k + 1 >= steps) break;
    const double bk = k > 0 ? be[k] : 0.0;
    for (int j = 0; j < n; ++j) nxt[j] = x[j] * cur[j] - a * cur[j] - bk * prev[j];
    double b2 = 0;
    for (int j = 0; j < n; ++j) b2 += w[j] * nxt[j] * nxt[j];
    const double b = std::sqrt(b2);
    if (b < 1e-15) {
      be.push_back(0.0);
      break;
    }
    be.push_back(b);
    for (int j = 0; j < n; ++j) nxt[j] /= b;
    prev.swap(cur);
    cur.swap(nxt);
  }
  for (int k = 0; k < n_bath; ++k) {
    eps[k] = k < int(al.size()) ? al[k] : 0.0;
    v[k] = k < int(be.size()) ? be[k] : 0.0;
  }
}

std::vector<double> BroydenMixer::update(const std::vector<double>& x, const std::vector<double>& f, double alpha) {
  const int dim = int(x.size());
  if (int(f.size()) != dim) throw ArgError("broyden: input and proposal differ in length");
  std::vector<double> res(static_cast<size_t>(dim)), out(static_cast<size_t>(dim));
  for (int i = 0; i < dim; ++i) res[i] = f[i] - x[i];
  if (!have_prev_ || dim_ != dim || n_updates_ >= max_history_) {
    dim_ = dim;
    n_updates_ = 0;
    jinv_.assign(size_t(dim) * dim, 0.0);
    for (int i = 0; i < dim; ++i) jinv_[size_t(i) * dim + i] = -alpha;
    prev_x_ = x;
    prev_r_ = res;
    have_prev_ = true;
    for (int i = 0; i < dim; ++i) out[i] = (1.0 - alpha) * x[i] + alpha * f[i];
    return out;
  }
  const size_t ud = static_cast<size_t>(dim);
  std::vector<double> dx(ud), df(ud), jdf(ud), num(ud), dxj(ud);
  for (int i = 0; i < dim; ++i) {
    dx[i] = x[i] - prev_x_[i];
    df[i] = res[i] - prev_r_[i];
  }
  for (int i = 0; i < dim; ++i) {
    double s = 0;
    for (int j = 0; j < dim; ++j) s += jinv_[size_t(i) * dim + j] * df[j];
    jdf[i] = s;
    num[i] = dx[i] - jdf[i];
  }
  for (int j = 0; j < dim; ++j) {
    double s = 0;
    for (int i = 0; i < dim; ++i) s += dx[i] * jinv_[size_t(i) * dim + j];
    dxj[j] = s;
  }
  double den = 0;
  for (int j = 0; j < dim; ++j) den += dxj[j] * df[j];
  if (std::fabs(den) > 1e-15)
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j) jinv_[size_t(i) * dim + j] += num[i] * dxj[j] / den;
  for (int i = 0; i < dim; ++i) {
    double c = 0;
    for (int j = 0; j < dim; ++j) c += jinv_[size_t(i) * dim + j] * res[j];
    out[i] = x[i] - c;
  }
  prev_x_ = x;
  prev_r_ = res;
  ++n_updates_;
  return out;
}

int krylov_tridiag(Context* ctx, int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                   const double* v0, int m, double* alpha, double* beta) {
  if (n < 1 || m < 1) throw ArgError("krylov: empty problem");
  DBuf<double> V(ctx, size_t(m + 1) * size_t(n)), w(ctx, size_t(n)), h(ctx, size_t(m + 1)),
      h2(ctx, size_t(m + 1)), nr(ctx, 1);
  double* Vp = V.get();
  TK_HIP(hipMemcpyAsync(Vp, v0, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  dot_many(ctx, Vp, n, 1, Vp, n, nr.get());
  double b0 = 0;
  nr.download(&b0, 1);
  ctx->sync();
  b0 = std::sqrt(b0);
  beta[0] = b0;
  if (!(b0 > 0)) return 0;
  scale(ctx, Vp, 1.0 / b0, n);
  double anorm = 0;
  int steps = 0;
  for (int j = 0; j < m; ++j) {
    double* q = Vp + int64_t(j) * n;
    hipLaunchKernelGGL(csr_spmv_kernel, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, n, rowptr, col, val, q,
                       w.get());
    TK_HIP(hipGetLastError());
    // two classical Gram-Schmidt passes against q_0..q_j
    dot_many(ctx, Vp, n, j + 1, w.get(), n, h.get());
    subtract_projection(ctx, Vp, n, j + 1, h.get(), w.get(), n);
    dot_many(ctx, Vp, n, j + 1, w.get(), n, h2.get());
    subtract_projection(ctx, Vp, n, j + 1, h2.get(), w.get(), n);
    dot_many(ctx, w.get(), n, 1, w.get(), n, nr.get());
    double a1 = 0, a2 = 0, b2 = 0;
    TK_HIP(hipMemcpyAsync(&a1, h.get() + j, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipMemcpyAsync(&a2, h2.get() + j, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    nr.download(&b2, 1);
    ctx->sync();
    alpha[j] = a1 + a2;
    steps = j + 1;
    const double b = std::sqrt(std::max(b2, 0.0));
    anorm = std::max(anorm, std::fabs(alpha[j]) + b + (j > 0 ? beta[j] : 0.0));
    if (j + 1 >= m) break;
    if (b <= 1e-13 * anorm) break;
    beta[j + 1] = b;
    double* qn = Vp + int64_t(j + 1) * n;
    TK_HIP(hipMemcpyAsync(qn, w.get(), size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    scale(ctx, qn, 1.0 / b, n);
  }
  ctx->sync();
  return steps;
}

// ---------------------------------------------------------------------------
Mlp::Mlp(Context* ctx, const std::vector<int>& dims, uint64_t seed) : ctx_(ctx), dims_(dims) {
  if (dims.size() < 2) throw ArgError("mlp: at least an input and an output layer");
  for (int d : dims)
    if (d < 1 || d > 65536) throw ArgError("mlp: layer width outside 1..65536");
  for (size_t l = 1; l < dims.size(); ++l) {
    w_off_.push_back(n_params_);
    n_params_ += int64_t(dims[l - 1]) * dims[l];
    b_off_.push_back(n_params_);
    n_params_ += dims[l];
  }
  // Glorot-uniform weights from a splitmix64 stream, zero biases
  std::vector<float> h(size_t(n_params_), 0.0f);
  uint64_t s = seed;
  auto next = [&s]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (size_t l = 1; l < dims.size(); ++l) {
    const double a = std::sqrt(6.0 / double(dims[l - 1] + dims[l]));
    for (int64_t e = 0; e < int64_t(dims[l - 1]) * dims[l]; ++e) h[w_off_[l - 1] + e] = float((2.0 * next() - 1.0) * a);
  }
  p_.reset(ctx, size_t(n_params_));
  g_.reset(ctx, size_t(n_params_));
  m_.reset(ctx, size_t(n_params_));
  v_.reset(ctx, size_t(n_params_));
  p_.upload(h.data(), h.size());
  std::vector<float> z(size_t(n_params_), 0.0f);
  g_.upload(z.data(), z.size());
  m_.upload(z.data(), z.size());
  v_.upload(z.data(), z.size());
  ctx_->sync();
}

void Mlp::get_params(float* host) const {
  p_.download(host, size_t(n_params_));
  ctx_->sync();
}
void Mlp::set_params(const float* host) {
  p_.upload(host, size_t(n_params_));
  ctx_->sync();
}

void Mlp::ensure(int batch) {
  if (batch < 1) throw ArgError("mlp: batch must be positive");
  if (batch <= batch_cap_) return;
  batch_cap_ = batch;
  act_off_.clear();
  int64_t tot = 0, dmax = 0;
  for (int d : dims_) {
    act_off_.push_back(tot);
    tot += int64_t(batch) * d;
    dmax = std::max<int64_t>(dmax, d);
  }
  act_.reset(ctx_, size_t(tot));
  dact_.reset(ctx_, size_t(2 * int64_t(batch) * dmax));
}

void Mlp::forward(int batch, const float* x, float* y) {
  ensure(batch);
  const int L = int(dims_.size()) - 1;
  float* act = act_.get();
  const int64_t an = int64_t(act_.size());
  TK_HIP(hipMemcpyAsync(act, x, size_t(batch) * dims_[0] * sizeof(float), hipMemcpyDeviceToDevice, ctx_->stream));
  for (int l = 1; l <= L; ++l) {
    const int din = dims_[l - 1], dout = dims_[l];
    gemm_f32(ctx_, 0, batch, dout, din, act, act_off_[l - 1], din, an, p_.get(), w_off_[l - 1], dout, n_params_, act,
             act_off_[l], dout, an);
    hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for(int64_t(batch) * dout)), dim3(256), 0, ctx_->stream,
                       int64_t(batch), dout, p_.get() + b_off_[l - 1], act + act_off_[l], l < L ? 1 : 0);
    TK_HIP(hipGetLastError());
  }
  if (y)
    TK_HIP(hipMemcpyAsync(y, act + act_off_[L], size_t(batch) * dims_[L] * sizeof(float), hipMemcpyDeviceToDevice,
                          ctx_->stream));
  ctx_->sync();
}

float Mlp::loss_grad(int batch, const float* x, const float* t, float* grad_host) {
  forward(batch, x, nullptr);
  const int L = int(dims_.size()) - 1;
  float* act = act_.get();
  const int64_t an = int64_t(act_.size());
  const int64_t nout = int64_t(batch) * dims_[L];
  const int64_t half = int64_t(dact_.size() / 2);
  float* G = dact_.get();         // gradient wrt the current layer's output
  float* Gp = G + half;           // and wrt the previous layer's
  const unsigned nb = blocks_for(nout);
  DBuf<float> part(ctx_, nb);
  hipLaunchKernelGGL(mse_grad_kernel, dim3(nb), dim3(256), 0, ctx_->stream, nout, 1.0f / float(nout),
                     act + act_off_[L], t, G, part.get());
  TK_HIP(hipGetLastError());
  std::vector<float> ph(nb);
  part.download(ph.data(), nb);
  for (int l = L; l >= 1; --l) {
    const int din = dims_[l - 1], dout = dims_[l];
    // dW = A_{l-1}^T G (din x dout), db = column sums of G
    gemm_f32(ctx_, kTransA, din, dout, batch, act, act_off_[l - 1], din, an, G, 0, dout, half, g_.get(),
             w_off_[l - 1], dout, n_params_);
    hipLaunchKernelGGL(colsum_kernel, dim3(blocks_for(dout)), dim3(256), 0, ctx_->stream, batch, dout, G,
                       g_.get() + b_off_[l - 1]);
    TK_HIP(hipGetLastError());
    if (l > 1) {
      // G_{l-1} = (G W^T) * (1 - A_{l-1}^2)
      gemm_f32(ctx_, kTransB, batch, din, dout, G, 0, dout, half, p_.get(), w_off_[l - 1], dout, n_params_, Gp, 0,
               din, half);
      hipLaunchKernelGGL(tanh_back_kernel, dim3(blocks_for(int64_t(batch) * din)), dim3(256), 0, ctx_->stream,
                         int64_t(batch) * din, act + act_off_[l - 1], Gp);
      TK_HIP(hipGetLastError());
      std::swap(G, Gp);
    }
  }
  if (grad_host) g_.download(grad_host, size_t(n_params_));
  ctx_->sync();
  double loss = 0;
  for (float v : ph) loss += v;
  return float(loss / double(nout));
}

void Mlp::adam_step(float lr, float beta1, float beta2, float eps) {
  ++t_;
  const float c1 = 1.0f - std::pow(beta1, float(t_)), c2 = 1.0f - std::pow(beta2, float(t_));
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n_params_)), dim3(256), 0, ctx_->stream, n_params_, lr, beta1,
                     beta2, eps, c1, c2, g_.get(), m_.get(), v_.get(), p_.get());
  TK_HIP(hipGetLastError());
  ctx_->sync();
}

}  // namespace tk
