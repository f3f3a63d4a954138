// Batched FP64 symmetric eigensolver: parallel cyclic Jacobi (Brent-Luk
// round-robin ordering, n/2 disjoint rotations per round), one workgroup per
// matrix.  Matrices up to kLdsMax live in LDS; larger ones are rotated in
// place in a global workspace.  Eigenvalues are returned sorted (descending
// or ascending) with eigenvectors as matching columns.
//
// Used for (a) the Gram matrices of the block-diagonal truncated SVD behind
// H_eff (tk-dmrg/src/truncation/mod.rs:56-127 -> faer thin_svd,
// tk-linalg/src/device/faer.rs:191-240) and (b) the projected matrix of the
// Davidson solver (symmetric_eigen_small, tk-dmrg/src/eigensolver/mod.rs:532-601).
// Rotation: tau = (a_qq - a_pp) / (2 a_pq), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1/sqrt(1+t^2), s = t c; columns (p, q) <- (c p - s q, s p + c q), then rows.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "eig.hpp"

namespace tk {
namespace {

constexpr int kThreads = 256;
constexpr int kLdsMax = 80;  // 2 * 80^2 * 8 B = 100 KiB of LDS

struct Rot {
  int p, q;
  double c, s;
};

// Round r of the circle method on m (even) indices: pair 0 = (m-1, r), pair k =
// ((r+k) mod (m-1), (r-k) mod (m-1)).
__device__ __forceinline__ void pair_of(int r, int k, int m, int& p, int& q) {
  if (k == 0) {
    p = r;
    q = m - 1;
  } else {
    p = (r + k) % (m - 1);
    q = (r - k + (m - 1)) % (m - 1);
  }
  if (p > q) {
    int t = p;
    p = q;
    q = t;
  }
}

template <bool kLds>
__device__ void jacobi(double* A, double* V, int n, int lda, Rot* rots, int* rotated, int max_sweeps) {
  const int tid = threadIdx.x;
  const int m = n + (n & 1);  // padded to even; index n (if padded) is a dummy
  for (int e = tid; e < n * n; e += kThreads) V[(e / n) * lda + e % n] = (e / n == e % n) ? 1.0 : 0.0;
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) *rotated = 0;
    __syncthreads();
    for (int r = 0; r < m - 1; ++r) {
      // rotation parameters for the m/2 pairs of this round
      for (int k = tid; k < m / 2; k += kThreads) {
        int p, q;
        pair_of(r, k, m, p, q);
        Rot ro{p, q, 1.0, 0.0};
        if (q < n) {
          const double apq = A[p * lda + q];
          const double app = A[p * lda + p], aqq = A[q * lda + q];
          if (fabs(apq) > 1e-300 && fabs(apq) > 1e-16 * sqrt(fabs(app) * fabs(aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            ro.c = 1.0 / sqrt(1.0 + t * t);
            ro.s = t * ro.c;
            *rotated = 1;
          }
        }
        rots[k] = ro;
      }
      __syncthreads();
      // A <- A J and V <- V J (columns p, q of every pair)
      for (int e = tid; e < (m / 2) * n; e += kThreads) {
        const int k = e / n, row = e - k * n;
        const Rot ro = rots[k];
        if (ro.q >= n || ro.s == 0.0) continue;
        double ap = A[row * lda + ro.p], aq = A[row * lda + ro.q];
        A[row * lda + ro.p] = ro.c * ap - ro.s * aq;
        A[row * lda + ro.q] = ro.s * ap + ro.c * aq;
        double vp = V[row * lda + ro.p], vq = V[row * lda + ro.q];
        V[row * lda + ro.p] = ro.c * vp - ro.s * vq;
        V[row * lda + ro.q] = ro.s * vp + ro.c * vq;
      }
      __syncthreads();
      // A <- J^T A (rows p, q of every pair)
      for (int e = tid; e < (m / 2) * n; e += kThreads) {
        const int k = e / n, col = e - k * n;
        const Rot ro = rots[k];
        if (ro.q >= n || ro.s == 0.0) continue;
        double ap = A[ro.p * lda + col], aq = A[ro.q * lda + col];
        A[ro.p * lda + col] = ro.c * ap - ro.s * aq;
        A[ro.q * lda + col] = ro.s * ap + ro.c * aq;
      }
      __syncthreads();
      for (int k = tid; k < m / 2; k += kThreads) {
        const Rot ro = rots[k];
        if (ro.q < n && ro.s != 0.0) {
          A[ro.p * lda + ro.q] = 0.0;
          A[ro.q * lda + ro.p] = 0.0;
        }
      }
      __syncthreads();
    }
    if (*rotated == 0) break;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void eig_kernel(const EigJob* __restrict__ jobs, int n_jobs,
                                                       double* __restrict__ Abuf, double* __restrict__ Vws,
                                                       double* __restrict__ W, double* __restrict__ Vout,
                                                       int descending, int max_sweeps) {
  __shared__ double lds[2 * kLdsMax * kLdsMax];
  __shared__ Rot rots[512];
  __shared__ int rotated;
  __shared__ double evals[1024];
  __shared__ int rank_of[1024];
  for (int j = blockIdx.x; j < n_jobs; j += gridDim.x) {
    const EigJob job = jobs[j];
    const int n = job.n;
    double* A;
    double* V;
    int lda;
    const bool in_lds = n <= kLdsMax;
    if (in_lds) {
      A = lds;
      V = lds + kLdsMax * kLdsMax;
      lda = n;
      for (int e = threadIdx.x; e < n * n; e += kThreads) A[e] = Abuf[job.a_off + e];
      __syncthreads();
      jacobi<true>(A, V, n, lda, rots, &rotated, max_sweeps);
    } else {
      A = Abuf + job.a_off;
      V = Vws + job.a_off;
      lda = n;
      jacobi<false>(A, V, n, lda, rots, &rotated, max_sweeps);
    }
    __syncthreads();
    // rank sort of the diagonal (ties broken by index)
    for (int i = threadIdx.x; i < n; i += kThreads) evals[i] = A[i * lda + i];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kThreads) {
      const double x = evals[i];
      int rk = 0;
      for (int k = 0; k < n; ++k) {
        const double y = evals[k];
        const bool before = descending ? (y > x || (y == x && k < i)) : (y < x || (y == x && k < i));
        rk += before ? 1 : 0;
      }
      rank_of[i] = rk;
      W[job.w_off + rk] = x;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += kThreads) {
      const int row = e / n, col = e - row * n;
      Vout[job.v_off + int64_t(row) * n + rank_of[col]] = V[row * lda + col];
    }
    __syncthreads();
  }
}

}  // namespace

void launch_eig(Context* ctx, const EigJob* jobs_dev, int n_jobs, double* A, double* Vws, double* W, double* Vout,
                bool descending, int max_sweeps) {
  if (n_jobs <= 0) return;
  int grid = std::min(n_jobs, ctx->n_cu * 4);
  hipLaunchKernelGGL(eig_kernel, dim3(grid), dim3(kThreads), 0, ctx->stream, jobs_dev, n_jobs, A, Vws, W, Vout,
                     descending ? 1 : 0, max_sweeps);
  TK_HIP(hipGetLastError());
}

int eig_max_n() { return 1024; }

}  // namespace tk
