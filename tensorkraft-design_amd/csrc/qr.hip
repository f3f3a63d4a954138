// Thin Householder QR on the device, blocked by panels of 32 columns.
//
// Reference: LinAlgBackend::qr (tk-linalg/src/traits.rs:148) implemented by
// faer's Householder QR with compute_thin_q / compute_thin_r
// (tk-linalg/src/device/faer.rs:292-324); callers are the MPS
// canonicalisation sweeps (tk-dmrg/src/mps/mod.rs:456, 646).
//
// Panel p (columns p0..p0+pb-1) is factored by one workgroup straight in the
// working copy: per column a reflector H = I - tau v v^T (v_0 = 1, beta =
// -sign(alpha) |x|, tau = 0 for an already-reduced column), one reduction of
// v against every other panel column (the columns after it are updated, the
// columns before it give the column of the compact-WY factor T, forward and
// columnwise: T[0:j, j] = -tau_j T[0:j, 0:j] V[:, 0:j]^T v_j).  The trailing
// columns are then updated as C <- C - V (T^T (V^T C)) by three grouped MFMA
// contractions.  Q is formed from [I; 0] by the panels in reverse order,
// Q <- Q - V (T (V^T Q)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "backend.hpp"
#include "kernels.hpp"

namespace tk {
namespace {

constexpr int kQrNb = 32;
constexpr int kQrThreads = 256;

__device__ double qr_block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int q = 0; q < kQrThreads / 64; ++q) s += red[q];
  return s;
}

// Factor the panel W[p0:m, p0:p0+pb) (row-major, leading dimension ld) in
// place; write V explicitly ((m - p0) x pb, unit diagonal, zeros above) and
// T (pb x pb, upper triangular).
__global__ __launch_bounds__(kQrThreads) void qr_panel_kernel(double* __restrict__ W, int64_t m, int64_t ld,
                                                              int64_t p0, int pb, double* __restrict__ V,
                                                              double* __restrict__ T) {
  __shared__ double red[kQrThreads / 64];
  __shared__ double dsum[kQrNb][kQrThreads / 64];
  __shared__ double d[kQrNb];
  __shared__ double Ts[kQrNb][kQrNb];
  __shared__ double z[kQrNb];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t mr = m - p0;
  double* P = W + p0 * ld + p0;  // panel origin
  for (int e = tid; e < kQrNb * kQrNb; e += kQrThreads) Ts[e / kQrNb][e % kQrNb] = 0.0;
  for (int j = 0; j < pb; ++j) {
    // x = P[j:, j]; sigma = |x[1:]|^2
    double loc = 0;
    for (int64_t r = j + 1 + tid; r < mr; r += kQrThreads) {
      const double x = P[r * ld + j];
      loc += x * x;
    }
    const double sigma = qr_block_sum(loc, red);
    const double alpha = P[int64_t(j) * ld + j];
    double tau = 0, beta = alpha, scale = 0;
    if (sigma > 0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    __syncthreads();  // every thread has read alpha before it is overwritten
    if (sigma > 0)
      for (int64_t r = j + 1 + tid; r < mr; r += kQrThreads) P[r * ld + j] *= scale;
    if (tid == 0) P[int64_t(j) * ld + j] = beta;
    __syncthreads();
    // d_c = sum_{r >= j} v_r P[r, c] for every other panel column (v_j = 1)
    double part[kQrNb];
#pragma unroll
    for (int c = 0; c < kQrNb; ++c) part[c] = 0;
    for (int64_t r = j + tid; r < mr; r += kQrThreads) {
      const double vr = r == j ? 1.0 : P[r * ld + j];
      const double* row = P + r * ld;
#pragma unroll
      for (int c = 0; c < kQrNb; ++c)
        if (c < pb && c != j) part[c] += vr * row[c];
    }
#pragma unroll
    for (int c = 0; c < kQrNb; ++c) {
      double v = part[c];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) dsum[c][wv] = v;
    }
    __syncthreads();
    if (tid < kQrNb) {
      double s = 0;
#pragma unroll
      for (int q = 0; q < kQrThreads / 64; ++q) s += dsum[tid][q];
      d[tid] = s;
    }
    __syncthreads();
    // columns after j: P[r, c] -= tau v_r d_c
    if (tau != 0)
      for (int64_t r = j + tid; r < mr; r += kQrThreads) {
        const double vr = r == j ? 1.0 : P[r * ld + j];
        double* row = P + r * ld;
        for (int c = j + 1; c < pb; ++c) row[c] -= tau * vr * d[c];
      }
    // T column j from z = V[:, 0:j]^T v_j = d[0:j]
    if (tid == 0) {
      for (int c = 0; c < j; ++c) z[c] = d[c];
      for (int i = 0; i < j; ++i) {
        double s = 0;
        for (int c = i; c < j; ++c) s += Ts[i][c] * z[c];
        Ts[i][j] = -tau * s;
      }
      Ts[j][j] = tau;
    }
    __syncthreads();
  }
  // explicit V and T
  for (int64_t e = tid; e < mr * pb; e += kQrThreads) {
    const int64_t r = e / pb;
    const int c = int(e - r * pb);
    V[e] = r == c ? 1.0 : (r > c ? P[r * ld + c] : 0.0);
  }
  for (int e = tid; e < pb * pb; e += kQrThreads) T[e] = Ts[e / pb][e % pb];
}

// R (k x n) = upper triangle of W[0:k, :]; Q (m x k) = [I_k; 0]
__global__ void qr_init_kernel(const double* __restrict__ W, int64_t m, int64_t n, int64_t k, double* __restrict__ R,
                               double* __restrict__ Q) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < k * n; e += stride) {
    const int64_t i = e / n, j = e - i * n;
    R[e] = j >= i ? W[i * n + j] : 0.0;
  }
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m * k; e += stride) {
    const int64_t i = e / k, j = e - i * k;
    Q[e] = i == j ? 1.0 : 0.0;
  }
}

}  // namespace

void qr(Context* ctx, int64_t m, int64_t n, const double* A, double* Q, double* R) {
  if (m <= 0 || n <= 0) throw ArgError("qr: matrix must have nonzero dimensions");
  if (!A || !Q || !R) throw ArgError("qr: null pointer");
  const int64_t k = std::min(m, n);
  DBuf<double> W(ctx, size_t(m * n));
  TK_HIP(hipMemcpyAsync(W.get(), A, size_t(m * n) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  // per panel: V ((m - p0) x pb) and T (pb x pb)
  std::vector<int64_t> voff, toff, p0s;
  std::vector<int> pbs;
  int64_t vtot = 0, ttot = 0;
  for (int64_t p0 = 0; p0 < k; p0 += kQrNb) {
    const int pb = int(std::min<int64_t>(kQrNb, k - p0));
    p0s.push_back(p0);
    pbs.push_back(pb);
    voff.push_back(vtot);
    toff.push_back(ttot);
    vtot += (m - p0) * pb;
    ttot += int64_t(pb) * pb;
  }
  DBuf<double> Vb(ctx, size_t(vtot)), Tb(ctx, size_t(ttot));
  DBuf<double> Y(ctx, size_t(kQrNb * std::max(n, k))), Y2(ctx, size_t(kQrNb * std::max(n, k)));
  for (size_t p = 0; p < p0s.size(); ++p) {
    const int64_t p0 = p0s[p], mr = m - p0;
    const int pb = pbs[p];
    hipLaunchKernelGGL(qr_panel_kernel, dim3(1), dim3(kQrThreads), 0, ctx->stream, W.get(), m, n, p0, pb,
                       Vb.get() + voff[p], Tb.get() + toff[p]);
    TK_HIP(hipGetLastError());
    const int64_t nc = n - p0 - pb;
    if (nc <= 0) continue;
    double* C = W.get() + p0 * n + p0 + pb;
    const double* Vp = Vb.get() + voff[p];
    const double* Tp = Tb.get() + toff[p];
    // C <- (I - V T^T V^T) C
    dense_gemm(ctx, pb, nc, mr, 1.0, Vp, 1, pb, C, n, 1, 0.0, Y.get(), nc, 1);
    dense_gemm(ctx, pb, nc, pb, 1.0, Tp, 1, pb, Y.get(), nc, 1, 0.0, Y2.get(), nc, 1);
    dense_gemm(ctx, mr, nc, pb, -1.0, Vp, pb, 1, Y2.get(), nc, 1, 1.0, C, n, 1);
  }
  const int grid = int(std::max<int64_t>(1, std::min<int64_t>((std::max(k * n, m * k) + 255) / 256,
                                                              int64_t(ctx->n_cu) * 8)));
  hipLaunchKernelGGL(qr_init_kernel, dim3(grid), dim3(256), 0, ctx->stream, W.get(), m, n, k, R, Q);
  TK_HIP(hipGetLastError());
  // Q = H_1 ... H_k [I; 0], panels applied last to first
  for (size_t pp = p0s.size(); pp-- > 0;) {
    const int64_t p0 = p0s[pp], mr = m - p0;
    const int pb = pbs[pp];
    double* Qs = Q + p0 * k;
    const double* Vp = Vb.get() + voff[pp];
    const double* Tp = Tb.get() + toff[pp];
    dense_gemm(ctx, pb, k, mr, 1.0, Vp, 1, pb, Qs, k, 1, 0.0, Y.get(), k, 1);
    dense_gemm(ctx, pb, k, pb, 1.0, Tp, pb, 1, Y.get(), k, 1, 0.0, Y2.get(), k, 1);
    dense_gemm(ctx, mr, k, pb, -1.0, Vp, pb, 1, Y2.get(), k, 1, 1.0, Qs, k, 1);
  }
  ctx->sync();
}

}  // namespace tk
