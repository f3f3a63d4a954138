// FP64 Krylov-vector kernels for the device-resident Davidson/Lanczos and the
// precision conversions of the mixed matvec.  All reductions are two-pass with
// a fixed summation order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "wave.hpp"

namespace tk {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxVec = 128;


// partial[i * nchunks + c] = sum over chunk c of V_i . x for the NV vectors
// i = i0 + NV blockIdx.y + q of a launch; chunks of kDotChunk elements, one
// barrier.  The loads carry no branch (an element index beyond n reads
// element n - 1 and meets x = 0), so each thread's loads are in flight
// together.
constexpr int kDotPer = 4;
constexpr int kDotChunk = kThreads * kDotPer;
constexpr int kVecGroup = 8;

// `active` (device flag, may be null): when it reads 0 the launch does nothing.
template <int NV>
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const double* __restrict__ V, int64_t ldv, int i0,
                                                               const double* __restrict__ x, int64_t n,
                                                               double* __restrict__ partial, int nchunks,
                                                               const int* __restrict__ active) {
  if (active != nullptr && *active == 0) return;
  __shared__ double red[kThreads / 64][NV];
  const int64_t base = int64_t(blockIdx.x) * kDotChunk + threadIdx.x;
  const int ib = i0 + int(blockIdx.y) * NV;
  double xr[kDotPer];
  int64_t at[kDotPer];
#pragma unroll
  for (int j = 0; j < kDotPer; ++j) {
    const int64_t idx = base + int64_t(j) * kThreads;
    at[j] = idx < n ? idx : n - 1;
    xr[j] = idx < n ? x[idx] : 0.0;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double* __restrict__ v = V + int64_t(ib + q) * ldv;
    s[q] = 0.0;
#pragma unroll
    for (int j = 0; j < kDotPer; ++j) s[q] += v[at[j]] * xr[j];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double t = wave_sum(s[q]);
    if (lane == 0) red[wave][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += red[w][threadIdx.x];
    partial[int64_t(ib + threadIdx.x) * nchunks + blockIdx.x] = t;
  }
}

// one launch of the last nv < kVecGroup vectors, from vector i0
template <int NV>
void launch_dot_rest(int nv, int nchunks, hipStream_t st, const double* V, int64_t ldv, int i0, const double* x,
                     int64_t n, double* part, const int* active) {
  if (nv == NV)
    hipLaunchKernelGGL(dot_partial_kernel<NV>, dim3(nchunks, 1), dim3(kThreads), 0, st, V, ldv, i0, x, n, part,
                       nchunks, active);
  else if constexpr (NV + 1 < kVecGroup)
    launch_dot_rest<NV + 1>(nv, nchunks, st, V, ldv, i0, x, n, part, active);
}

__global__ __launch_bounds__(kThreads) void reduce_rows_kernel(const double* __restrict__ partial, int nchunks,
                                                               double* __restrict__ out,
                                                               const int* __restrict__ active) {
  if (active != nullptr && *active == 0) return;
  __shared__ double red[kThreads / 64];
  const double* row = partial + int64_t(blockIdx.x) * nchunks;
  double s = 0.0;
  for (int c = threadIdx.x; c < nchunks; c += kThreads) s += row[c];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    out[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kThreads) void combine_kernel(const double* __restrict__ V, int64_t ldv, int nvec,
                                                           const double* __restrict__ c, double beta, double sign,
                                                           double* __restrict__ y, int64_t n,
                                                           const int* __restrict__ active) {
  if (active != nullptr && *active == 0) return;
  __shared__ double cs[kMaxVec];
  for (int i = threadIdx.x; i < nvec; i += kThreads) cs[i] = c[i];
  __syncthreads();
  for (int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x; j < n; j += int64_t(gridDim.x) * kThreads) {
    double s = 0.0;
    for (int i = 0; i < nvec; ++i) s += cs[i] * V[int64_t(i) * ldv + j];
    y[j] = (beta == 0.0 ? 0.0 : beta * y[j]) + sign * s;
  }
}

// One pass over the basis V and its image AV (nvec vectors each): the Ritz
// residual r = AV c - theta V c, written to t preconditioned as
// r / (theta - d) where |theta - d| > 1e-14 (d null: r itself), with per-block
// partial sums of r^2 (fixed grid-stride order) for its norm.
__global__ __launch_bounds__(kThreads) void ritz_residual_kernel(const double* __restrict__ V,
                                                                 const double* __restrict__ AV, int64_t ldv,
                                                                 int nvec, const double* __restrict__ c,
                                                                 const double* __restrict__ theta,
                                                                 const double* __restrict__ d,
                                                                 double* __restrict__ t, int64_t n,
                                                                 double* __restrict__ partial) {
  __shared__ double cs[kMaxVec];
  __shared__ double red[kThreads / 64];
  for (int i = threadIdx.x; i < nvec; i += kThreads) cs[i] = c[i];
  __syncthreads();
  const double th = *theta;
  double rr = 0.0;
  for (int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x; j < n; j += int64_t(gridDim.x) * kThreads) {
    double y = 0.0, ay = 0.0;
    for (int i = 0; i < nvec; ++i) {
      y += cs[i] * V[int64_t(i) * ldv + j];
      ay += cs[i] * AV[int64_t(i) * ldv + j];
    }
    const double r = ay - th * y;
    rr += r * r;
    if (d != nullptr) {
      const double den = th - d[j];
      t[j] = fabs(den) > 1e-14 ? r / den : r;
    } else {
      t[j] = r;
    }
  }
  rr = wave_sum(rr);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rr;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w];
    partial[blockIdx.x] = s;
  }
}

// The pass of ritz_residual_kernel for NV <= kFusedVec basis vectors,
// summing on the way, with the basis values held in registers, the first
// Gram-Schmidt pass's products of the t it writes: partial rows i < NV:
// V_i . t, row NV: t . t, row NV + 1: r . r (grid columns each).  t and
// r . r are those of ritz_residual_kernel (same order of every sum).
constexpr int kFusedVec = 16;

template <int NV>
__global__ __launch_bounds__(kThreads) void ritz_residual_dots_kernel(
    const double* __restrict__ V, const double* __restrict__ AV, int64_t ldv, const double* __restrict__ c,
    const double* __restrict__ theta, const double* __restrict__ d, double* __restrict__ t, int64_t n,
    double* __restrict__ partial, int grid) {
  constexpr int NQ = NV + 2;
  __shared__ double cs[NV];
  __shared__ double red[kThreads / 64][NQ];
  if (threadIdx.x < NV) cs[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  const double th = *theta;
  double pd[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) pd[q] = 0.0;
  for (int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x; j < n; j += int64_t(gridDim.x) * kThreads) {
    double vv[NV], av[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      vv[i] = V[int64_t(i) * ldv + j];
      av[i] = AV[int64_t(i) * ldv + j];
    }
    double y = 0.0, ay = 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      y += cs[i] * vv[i];
      ay += cs[i] * av[i];
    }
    const double r = ay - th * y;
    pd[NV + 1] += r * r;
    double tj = r;
    if (d != nullptr) {
      const double den = th - d[j];
      tj = fabs(den) > 1e-14 ? r / den : r;
    }
    t[j] = tj;
#pragma unroll
    for (int i = 0; i < NV; ++i) pd[i] += vv[i] * tj;
    pd[NV] += tj * tj;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double x = wave_sum(pd[q]);
    if (lane == 0) red[wave][q] = x;
  }
  __syncthreads();
  const int q = threadIdx.x;
  if (q < NQ) {
    double x = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) x += red[w][q];
    partial[int64_t(q) * grid + blockIdx.x] = x;
  }
}

// y <- y - sum_i h_i V_i for NV <= kFusedVec vectors (the arithmetic of
// combine_kernel with beta = 1, sign = -1), summing on the way the partial
// rows i < NV: V_i . y_new (when `dots`) and row NV: y_new . y_new.
// `active` as in dot_partial_kernel.
template <int NV>
__global__ __launch_bounds__(kThreads) void project_out_dots_kernel(const double* __restrict__ V, int64_t ldv,
                                                                    const double* __restrict__ h,
                                                                    double* __restrict__ y, int64_t n, int dots,
                                                                    double* __restrict__ partial, int grid,
                                                                    const int* __restrict__ active) {
  if (active != nullptr && *active == 0) return;
  constexpr int NQ = NV + 1;
  __shared__ double hs[NV];
  __shared__ double red[kThreads / 64][NQ];
  if (threadIdx.x < NV) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double pd[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) pd[q] = 0.0;
  for (int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x; j < n; j += int64_t(gridDim.x) * kThreads) {
    double vv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) vv[i] = V[int64_t(i) * ldv + j];
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) s += hs[i] * vv[i];
    const double yn = y[j] - s;
    y[j] = yn;
#pragma unroll
    for (int i = 0; i < NV; ++i) pd[i] += vv[i] * yn;
    pd[NV] += yn * yn;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double x = wave_sum(pd[q]);
    if (lane == 0) red[wave][q] = x;
  }
  __syncthreads();
  const int q = threadIdx.x;
  if ((dots && q < NV) || q == NV) {
    double x = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) x += red[w][q];
    partial[int64_t(q) * grid + blockIdx.x] = x;
  }
}

// the launch of a fused pass for nvec = NV .. kFusedVec vectors
template <int NV>
void launch_ritz_dots(int nvec, dim3 grid, hipStream_t st, const double* V, const double* AV, int64_t ldv,
                      const double* c, const double* theta, const double* d, double* t, int64_t n, double* part) {
  if (nvec == NV)
    hipLaunchKernelGGL(ritz_residual_dots_kernel<NV>, grid, dim3(kThreads), 0, st, V, AV, ldv, c, theta, d, t, n,
                       part, int(grid.x));
  else if constexpr (NV < kFusedVec)
    launch_ritz_dots<NV + 1>(nvec, grid, st, V, AV, ldv, c, theta, d, t, n, part);
}

template <int NV>
void launch_project_dots(int nvec, dim3 grid, hipStream_t st, const double* V, int64_t ldv, const double* h,
                         double* y, int64_t n, int dots, double* part, const int* active) {
  if (nvec == NV)
    hipLaunchKernelGGL(project_out_dots_kernel<NV>, grid, dim3(kThreads), 0, st, V, ldv, h, y, n, dots, part,
                       int(grid.x), active);
  else if constexpr (NV < kFusedVec)
    launch_project_dots<NV + 1>(nvec, grid, st, V, ldv, h, y, n, dots, part, active);
}

// Second Gram-Schmidt pass needed?  h[0..nvec) = V^T s and h[nvec] = s.s
// before the first pass; |s - V h|^2 = s.s - |h|^2 for orthonormal V.  The
// pass is repeated when the first one removed more than half of |s|^2.
__global__ void reorth_flag_kernel(const double* __restrict__ h, int nvec, int* __restrict__ flag) {
  if (threadIdx.x == 0) {
    double hh = 0.0;
    for (int i = 0; i < nvec; ++i) hh += h[i] * h[i];
    const double ss = h[nvec];
    flag[0] = (ss - hh < 0.5 * ss) ? 1 : 0;
  }
}

__global__ void scale_inv_sqrt_kernel(double* __restrict__ x, const double* __restrict__ nrm2, int64_t n) {
  const double s = *nrm2;
  const double f = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    x[j] *= f;
}

__global__ void scale_kernel(double* __restrict__ x, double a, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    x[j] *= a;
}
__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    y[j] += a * x[j];
}
__global__ void residual_kernel(const double* __restrict__ ay, const double* __restrict__ y, double theta,
                                double* __restrict__ r, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    r[j] = ay[j] - theta * y[j];
}
__global__ void precond_kernel(const double* __restrict__ r, const double* __restrict__ d, double theta,
                               double* __restrict__ t, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x) {
    double den = theta - d[j];
    t[j] = fabs(den) > 1e-14 ? r[j] / den : r[j];
  }
}
__global__ void f64_to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    y[j] = float(x[j]);
}
__global__ void f32_to_f64_kernel(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    y[j] = double(x[j]);
}
__global__ void fill_kernel(double* __restrict__ x, double v, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    x[j] = v;
}

int grid_for(Context* ctx, int64_t n) {
  int64_t g = (n + kThreads - 1) / kThreads;
  return int(std::max<int64_t>(1, std::min<int64_t>(g, int64_t(ctx->n_cu) * 16)));
}

double* scratch(Context* ctx, size_t n) { return static_cast<double*>(ctx->scratch(n * sizeof(double))); }

}  // namespace

void dot_many(Context* ctx, const double* V, int64_t ldv, int nvec, const double* x, int64_t n, double* out,
              const int* active) {
  if (nvec <= 0) return;
  if (nvec > kMaxVec) throw ArgError("dot_many: too many vectors");
  const int nchunks = int(std::max<int64_t>(1, (n + kDotChunk - 1) / kDotChunk));
  double* part = scratch(ctx, size_t(nchunks) * nvec);
  const int full = nvec / kVecGroup, rest = nvec % kVecGroup;
  if (full > 0)
    hipLaunchKernelGGL(dot_partial_kernel<kVecGroup>, dim3(nchunks, full), dim3(kThreads), 0, ctx->stream, V, ldv, 0,
                       x, n, part, nchunks, active);
  if (rest > 0) launch_dot_rest<1>(rest, nchunks, ctx->stream, V, ldv, full * kVecGroup, x, n, part, active);
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(nvec), dim3(kThreads), 0, ctx->stream, part, nchunks, out, active);
  TK_HIP(hipGetLastError());
}

void combine(Context* ctx, const double* V, int64_t ldv, int nvec, const double* c, double beta, double* y,
             int64_t n) {
  if (nvec > kMaxVec) throw ArgError("combine: too many vectors");
  hipLaunchKernelGGL(combine_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, V, ldv, nvec, c, beta,
                     1.0, y, n, static_cast<const int*>(nullptr));
  TK_HIP(hipGetLastError());
}

void subtract_projection(Context* ctx, const double* V, int64_t ldv, int nvec, const double* h, double* y,
                         int64_t n, const int* active) {
  if (nvec > kMaxVec) throw ArgError("subtract_projection: too many vectors");
  hipLaunchKernelGGL(combine_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, V, ldv, nvec, h, 1.0,
                     -1.0, y, n, active);
  TK_HIP(hipGetLastError());
}

void ritz_residual(Context* ctx, const double* V, const double* AV, int64_t ldv, int nvec, const double* c,
                   const double* theta, const double* diag, double* t, int64_t n, double* rn2) {
  if (nvec <= 0 || nvec > kMaxVec) throw ArgError("ritz_residual: vector count outside 1..128");
  const int grid = grid_for(ctx, n);
  double* part = scratch(ctx, size_t(grid));
  hipLaunchKernelGGL(ritz_residual_kernel, dim3(grid), dim3(kThreads), 0, ctx->stream, V, AV, ldv, nvec, c, theta,
                     diag, t, n, part);
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, part, grid, rn2,
                     static_cast<const int*>(nullptr));
  TK_HIP(hipGetLastError());
}

int fused_vec_max() { return kFusedVec; }

// Grid of the fused passes: one resident round (two workgroups per CU at
// their register use), so that the per-workgroup reduction of their 17 or 18
// sums is paid once per CU pair rather than once per 256 elements.
int fused_grid(Context* ctx, int64_t n) {
  const int64_t g = (n + kThreads - 1) / kThreads;
  return int(std::max<int64_t>(1, std::min<int64_t>(g, int64_t(ctx->n_cu) * 2)));
}

void ritz_residual_dots(Context* ctx, const double* V, const double* AV, int64_t ldv, int nvec, const double* c,
                        const double* theta, const double* diag, double* t, int64_t n, double* rn2, double* h) {
  if (nvec <= 0 || nvec > kFusedVec) throw ArgError("ritz_residual_dots: vector count outside 1..16");
  const int grid = fused_grid(ctx, n);
  double* part = scratch(ctx, size_t(grid) * (nvec + 2));
  launch_ritz_dots<1>(nvec, dim3(grid), ctx->stream, V, AV, ldv, c, theta, diag, t, n, part);
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(nvec + 1), dim3(kThreads), 0, ctx->stream, part, grid, h,
                     static_cast<const int*>(nullptr));
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, part + int64_t(nvec + 1) * grid,
                     grid, rn2, static_cast<const int*>(nullptr));
  TK_HIP(hipGetLastError());
}

void project_out_dots(Context* ctx, const double* V, int64_t ldv, int nvec, const double* h, double* y, int64_t n,
                      double* h_new, double* nrm2, const int* active) {
  if (nvec <= 0 || nvec > kFusedVec) throw ArgError("project_out_dots: vector count outside 1..16");
  const int grid = fused_grid(ctx, n);
  double* part = scratch(ctx, size_t(grid) * (nvec + 1));
  launch_project_dots<1>(nvec, dim3(grid), ctx->stream, V, ldv, h, y, n, h_new != nullptr ? 1 : 0, part, active);
  if (h_new != nullptr)
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(nvec), dim3(kThreads), 0, ctx->stream, part, grid, h_new, active);
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, part + int64_t(nvec) * grid, grid,
                     nrm2, active);
  TK_HIP(hipGetLastError());
}

void reorth_flag(Context* ctx, const double* h, int nvec, int* flag) {
  hipLaunchKernelGGL(reorth_flag_kernel, dim3(1), dim3(64), 0, ctx->stream, h, nvec, flag);
  TK_HIP(hipGetLastError());
}

void scale_by_inv_sqrt(Context* ctx, double* x, const double* nrm2, int64_t n) {
  hipLaunchKernelGGL(scale_inv_sqrt_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, x, nrm2, n);
  TK_HIP(hipGetLastError());
}
void scale(Context* ctx, double* x, double a, int64_t n) {
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, x, a, n);
  TK_HIP(hipGetLastError());
}
void axpy(Context* ctx, double a, const double* x, double* y, int64_t n) {
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, a, x, y, n);
  TK_HIP(hipGetLastError());
}
void residual(Context* ctx, const double* ay, const double* y, double theta, double* r, int64_t n) {
  hipLaunchKernelGGL(residual_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, ay, y, theta, r, n);
  TK_HIP(hipGetLastError());
}
void precondition(Context* ctx, const double* r, const double* diag, double theta, double* t, int64_t n) {
  hipLaunchKernelGGL(precond_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, r, diag, theta, t, n);
  TK_HIP(hipGetLastError());
}
void convert_f64_to_f32(Context* ctx, const double* x, float* y, int64_t n) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, x, y, n);
  TK_HIP(hipGetLastError());
}
KernelNode convert_f64_to_f32_node(Context* ctx, const double* x, float* y, int64_t n) {
  KernelNode k;
  k.func = reinterpret_cast<const void*>(&f64_to_f32_kernel);
  k.grid = dim3(grid_for(ctx, n));
  k.block = dim3(kThreads);
  k.nargs = 3;
  k.set(0, x);
  k.set(1, y);
  k.set(2, n);
  k.bind();
  return k;
}
void convert_f32_to_f64(Context* ctx, const float* x, double* y, int64_t n) {
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, x, y, n);
  TK_HIP(hipGetLastError());
}
void fill_f64(Context* ctx, double* x, double v, int64_t n) {
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, x, v, n);
  TK_HIP(hipGetLastError());
}

}  // namespace tk
