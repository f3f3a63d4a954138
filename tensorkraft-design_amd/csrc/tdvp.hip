// Krylov exponential on the device (see tdvp.hpp).
#include "tdvp.hpp"

#include <cmath>
#include <limits>

#include "kernels.hpp"

namespace tk {
namespace {

constexpr int kExpThreads = 256;
constexpr int kExpPer = (kExpKrylovMax * kExpKrylovMax + kExpThreads - 1) / kExpThreads;
static_assert(kExpKrylovMax <= 64, "the vector series keeps one row per lane of one wavefront");

// Deterministic sum over the workgroup (tree in LDS); every thread gets it.
__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int w = kExpThreads / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// c = exp(alpha H[:m, :m]) (beta e_1), H row-major with leading dimension ldh
// (mat_exp_times_vec, tdvp/mod.rs:278-382).  One workgroup.
__global__ __launch_bounds__(kExpThreads) void krylov_expm_kernel(const double* __restrict__ H, int ldh, int m,
                                                                  double alpha, double beta,
                                                                  double* __restrict__ c) {
  __shared__ double a[kExpKrylovMax * kExpKrylovMax];
  __shared__ double e[kExpKrylovMax * kExpKrylovMax];
  __shared__ double t[kExpKrylovMax * kExpKrylovMax];
  __shared__ double red[kExpThreads];
  const int tid = threadIdx.x, mm = m * m;
  double p = 0;
  for (int idx = tid; idx < mm; idx += kExpThreads) {
    const int i = idx / m, j = idx - i * m;
    a[idx] = alpha * H[int64_t(i) * ldh + j];
    p += fabs(a[idx]);
  }
  const double norm = block_sum(p, red) / m;  // the reference's mean absolute row sum
  if (!(norm <= 1e300)) {
    for (int i = tid; i < m; i += kExpThreads) c[i] = __builtin_nan("");
    return;
  }
  const int s = norm > 0 ? max(0, int(ceil(log2(norm / 0.5)))) : 0;
  const double scale = ldexp(1.0, -s);
  for (int idx = tid; idx < mm; idx += kExpThreads) a[idx] *= scale;
  __syncthreads();
  if (s == 0) {
    // vector Taylor series in wavefront 0, lane i holding row i (m <= 64):
    // the term's entries are broadcast by shuffles, its norm by a butterfly,
    // so the loop needs no workgroup barrier
    if (tid >= 64) return;
    double ti = tid == 0 ? beta : 0.0;  // current term
    double ei = ti;                     // partial sum
    for (int k = 1; k <= 20; ++k) {
      double nt = 0;
      for (int j = 0; j < m; ++j) {
        const double tj = __shfl(ti, j, 64);
        nt += (tid < m ? a[tid * m + j] : 0.0) * tj;
      }
      nt = tid < m ? nt / k : 0.0;
      ti = nt;
      ei += nt;
      double sq = nt * nt;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
      if (sqrt(sq) < 1e-16 * fabs(beta)) break;
    }
    if (tid < m) c[tid] = ei;
    return;
  }
  // matrix Taylor series of a 2^-s, then s squarings
  for (int idx = tid; idx < mm; idx += kExpThreads) {
    const int i = idx / m, j = idx - i * m;
    e[idx] = t[idx] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  double r[kExpPer];
  for (int k = 1; k <= 20; ++k) {
    double sq = 0;
#pragma unroll
    for (int h = 0; h < kExpPer; ++h) {
      const int idx = tid + kExpThreads * h;
      r[h] = 0;
      if (idx < mm) {
        const int i = idx / m, j = idx - i * m;
        double acc = 0;
        for (int l = 0; l < m; ++l) acc += a[i * m + l] * t[l * m + j];
        r[h] = acc / k;
        sq += r[h] * r[h];
      }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kExpPer; ++h) {
      const int idx = tid + kExpThreads * h;
      if (idx < mm) {
        t[idx] = r[h];
        e[idx] += r[h];
      }
    }
    const double tn = sqrt(block_sum(sq, red));
    if (tn < 1e-16) break;
  }
  for (int q = 0; q < s; ++q) {
#pragma unroll
    for (int h = 0; h < kExpPer; ++h) {
      const int idx = tid + kExpThreads * h;
      r[h] = 0;
      if (idx < mm) {
        const int i = idx / m, j = idx - i * m;
        double acc = 0;
        for (int l = 0; l < m; ++l) acc += e[i * m + l] * e[l * m + j];
        r[h] = acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kExpPer; ++h) {
      const int idx = tid + kExpThreads * h;
      if (idx < mm) e[idx] = r[h];
    }
    __syncthreads();
  }
  for (int i = tid; i < m; i += kExpThreads) c[i] = e[i * m] * beta;
}

// H[i * ldh + j] (+)= h[i] for i < count
__global__ void hcol_kernel(double* __restrict__ H, int ldh, int j, const double* __restrict__ h, int count,
                            int accumulate) {
  for (int i = threadIdx.x; i < count; i += blockDim.x)
    H[int64_t(i) * ldh + j] = (accumulate ? H[int64_t(i) * ldh + j] : 0.0) + h[i];
}

__global__ void sqrt_store_kernel(const double* __restrict__ nrm2, double* __restrict__ out) {
  if (threadIdx.x == 0) *out = sqrt(fmax(0.0, *nrm2));
}

struct PlanUser {
  HeffPlan* plan;
};
void plan_matvec_fp64(const double* x, double* y, int, void* user) {
  static_cast<PlanUser*>(user)->plan->apply_fp64(x, y);
}

}  // namespace

ExpKrylovResult exp_krylov_generic(Context* ctx, int64_t n, MatvecFn mv, void* user, const double* v, double alpha,
                                   int krylov_dim, double tol, double* out, double adaptive_tol) {
  ExpKrylovResult res;
  if (n < 0 || krylov_dim < 0) throw ArgError("exp_krylov: negative size");
  if (krylov_dim > kExpKrylovMax) throw ArgError("exp_krylov: krylov_dim above 48");
  hipStream_t s = ctx->stream;
  const int m = int(std::min<int64_t>(krylov_dim, n));
  if (m == 0) {
    if (n > 0 && out != v) TK_HIP(hipMemcpyAsync(out, v, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
    TK_HIP(hipStreamSynchronize(s));
    return res;
  }
  DBuf<double> Vb(ctx, size_t(m) * n), wb(ctx, n), hb(ctx, m), Hb(ctx, size_t(m + 1) * m), cb(ctx, m), sc(ctx, 2);
  double* V = Vb.get();
  double* w = wb.get();
  double* H = Hb.get();
  double* nrm2 = sc.get();
  fill_f64(ctx, H, 0.0, int64_t(m + 1) * m);
  dot_many(ctx, v, n, 1, v, n, nrm2);
  double host = 0;
  TK_HIP(hipMemcpyAsync(&host, nrm2, sizeof(double), hipMemcpyDeviceToHost, s));
  TK_HIP(hipStreamSynchronize(s));
  const double beta = std::sqrt(std::max(0.0, host));
  res.norm_in = beta;
  if (!std::isfinite(beta)) throw ArgError("exp_krylov: start vector is not finite");
  if (beta < 1e-30) {
    fill_f64(ctx, out, 0.0, n);
    TK_HIP(hipStreamSynchronize(s));
    return res;
  }
  TK_HIP(hipMemcpyAsync(V, v, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
  scale(ctx, V, 1.0 / beta, n);
  int used = m;
  const bool adaptive = adaptive_tol > 0;
  double c[kExpKrylovMax];  // host copy of the small exponential's result
  for (int j = 0; j < m; ++j) {
    const double* vj = V + int64_t(j) * n;
    mv(vj, w, 1, user);
    ++res.matvecs;
    // column j of H: two classical Gram-Schmidt passes, coefficients summed
    dot_many(ctx, V, n, j + 1, w, n, hb.get());
    hipLaunchKernelGGL(hcol_kernel, dim3(1), dim3(64), 0, s, H, m, j, hb.get(), j + 1, 0);
    subtract_projection(ctx, V, n, j + 1, hb.get(), w, n);
    dot_many(ctx, V, n, j + 1, w, n, hb.get());
    hipLaunchKernelGGL(hcol_kernel, dim3(1), dim3(64), 0, s, H, m, j, hb.get(), j + 1, 1);
    subtract_projection(ctx, V, n, j + 1, hb.get(), w, n);
    dot_many(ctx, w, n, 1, w, n, nrm2);
    hipLaunchKernelGGL(sqrt_store_kernel, dim3(1), dim3(64), 0, s, nrm2, H + int64_t(j + 1) * m + j);
    if (adaptive) {
      hipLaunchKernelGGL(krylov_expm_kernel, dim3(1), dim3(kExpThreads), 0, s, H, m, j + 1, alpha, beta, cb.get());
      TK_HIP(hipMemcpyAsync(c, cb.get(), size_t(j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    TK_HIP(hipMemcpyAsync(&host, nrm2, sizeof(double), hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    TK_HIP(hipGetLastError());
    const double hn = std::sqrt(std::max(0.0, host));
    if (!std::isfinite(hn)) throw ArgError("exp_krylov: the matvec produced a non-finite vector");
    if (hn < tol) {  // happy breakdown: the Krylov space is invariant
      used = j + 1;
      break;
    }
    if (adaptive) {
      double cn = 0;
      for (int i = 0; i <= j; ++i) cn += c[i] * c[i];
      if (hn * std::fabs(c[j]) <= adaptive_tol * std::sqrt(cn)) {
        used = j + 1;
        break;
      }
    }
    if (j + 1 < m) {
      double* vn = V + int64_t(j + 1) * n;
      TK_HIP(hipMemcpyAsync(vn, w, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
      scale_by_inv_sqrt(ctx, vn, nrm2, n);
    }
  }
  // with the early stop on, cb already holds exp(alpha H_used) beta e_1
  if (!adaptive)
    hipLaunchKernelGGL(krylov_expm_kernel, dim3(1), dim3(kExpThreads), 0, s, H, m, used, alpha, beta, cb.get());
  combine(ctx, V, n, used, cb.get(), 0.0, out, n);
  TK_HIP(hipMemcpyAsync(c, cb.get(), size_t(used) * sizeof(double), hipMemcpyDeviceToHost, s));
  TK_HIP(hipStreamSynchronize(s));
  TK_HIP(hipGetLastError());
  double cn = 0;
  for (int i = 0; i < used; ++i) cn += c[i] * c[i];
  if (!std::isfinite(cn)) throw ArgError("exp_krylov: exp(alpha H) overflowed");
  res.krylov_dim = used;
  res.norm_out = std::sqrt(cn);
  return res;
}

ExpKrylovResult exp_krylov(Context* ctx, HeffPlan& plan, const double* v, double alpha, int krylov_dim, double tol,
                           double* out, double adaptive_tol) {
  PlanUser u{&plan};
  return exp_krylov_generic(ctx, plan.psi().total, plan_matvec_fp64, &u, v, alpha, krylov_dim, tol, out, adaptive_tol);
}

}  // namespace tk
