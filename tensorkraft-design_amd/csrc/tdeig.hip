// Batched FP64 symmetric eigensolver: Householder tridiagonalisation, Sturm
// bisection, inverse iteration, blocked Gram-Schmidt and back-transformation.
// Reference call sites and the method are summarised in tdeig.hpp.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>

#include "tdeig.hpp"
#include "tdcoop.hpp"
#include "wave.hpp"

#include "contract.hpp"
#include "kernels.hpp"

namespace tk {
namespace {

constexpr double kEps = DBL_EPSILON;
constexpr int kLuFields = 4;
// lanes per eigenvalue in the multisection (TK_BIS_LANES = 4, 8 or 16 overrides, for tuning)
int bis_lanes() {
  static const int l = [] {
    const char* e = std::getenv("TK_BIS_LANES");
    const int v = e ? std::atoi(e) : 0;
    return (v == 4 || v == 8 || v == 16) ? v : 16;
  }();
  return l;
}

struct DevJob {
  int64_t a_off;   // matrix / reflectors (and output columns in Z)
  int64_t w_off;   // eigenvalues, d, e, tau
  int64_t z_off;   // inverse-iteration vectors: n x k row-major (ld k)
  int64_t lu_off;  // LU factors: kLuFields doubles per (row, vector), row-major over (row, vector)
  int32_t n, k;
  int32_t job;     // index of the job in the caller's list
  int32_t pad;
};


// Sum over the workgroup; `red` holds 2 * (NT / 64) doubles used alternately,
// so consecutive calls need no second barrier.
template <int NT>
__device__ double block_sum(double x, double* red, int& parity) {
  constexpr int NW = NT / 64;
  x = wave_sum(x);
  double* r = red + parity * NW;
  parity ^= 1;
  if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += r[i];
  return s;
}

// Number of eigenvalues of the tridiagonal (d, e^2) smaller than x.
// a / q through the hardware reciprocal and one Newton step (callers keep
// |q| >= DBL_MIN: the Sturm recurrence by pivmin, inverse iteration by its
// pivot guard and an IEEE division for a smaller off-diagonal): at most 18
// ulps from the IEEE quotient over the normal range and 15 with divisors
// near DBL_MIN (measured on 250,000 random operand pairs; 5 ulps for
// subnormal quotients), i.e. a relative perturbation of e_i^2 of 4e-15 per
// Sturm step, in a third of the IEEE division's instructions and dependent
// latency.  A second Newton step would bring it to about one ulp at a
// quarter more latency on the recurrence.
__device__ __forceinline__ double fast_quot(double e2, double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  return e2 * r;
}

// q[i] = fast_quot(a[i], b[i]): the quotient as the eigensolver's kernels form it.
__global__ void fast_quot_kernel(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ q,
                                 int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) q[i] = fast_quot(a[i], b[i]);
}

__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int cnt = q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - fast_quot(e2[i - 1], q);
    if (fabs(q) <= pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Deterministic start vectors in [-1, 1).
__device__ __forceinline__ double hash_unit(uint32_t a, uint32_t b) {
  uint32_t h = a * 0x9E3779B1u + (b + 0x632BE5ABu) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return double(h >> 8) * (2.0 / 16777216.0) - 1.0;
}

// ---------------------------------------------------------------------------
// 1. Tridiagonalisation.  Step k: x = A[k][k+1:] (row k of the updated upper
// triangle = column k below the diagonal), reflector H_k = I - tau v v^T with
// v[k+1] = 1, T[k+1][k] = beta; p = A22 v, w = tau p - tau^2/2 (p.v) v,
// A22 -= v w^T + w v^T.  The update of step k-1 is applied to row k first,
// then to rows k+1.. in the same pass that forms A22 v for step k.
// T = max number of 64-column tiles of a row (n <= 64 T).
// One matrix per workgroup: the trailing block of size m at a_off (leading
// dimension ld) of a matrix whose column steps before it were done by the
// multi-workgroup kernel; `pending` >= 0 names the VP buffer pair holding
// that kernel's last reflector v and A22 v, whose rank-2 update is still to
// be applied (-1: a whole matrix, nothing pending).
struct TailJob {
  int64_t a_off;   // top-left element of the trailing block
  int64_t w_off;   // d, e, tau (and the VP vectors) of its first row
  int32_t m, ld;
  int32_t pending, pad;
};

// With kLds the block is copied into LDS (m <= 128) and only the reflector
// rows are written back.
template <int NT, int T, bool kLds>
__global__ __launch_bounds__(NT) void tridiag_kernel(const TailJob* __restrict__ jobs, double* __restrict__ A,
                                                     double* __restrict__ Dg, double* __restrict__ Eg,
                                                     double* __restrict__ TAUg, const double* __restrict__ VP,
                                                     int64_t w_n) {
  constexpr int NW = NT / 64;
  constexpr int R = (64 * T + NT - 1) / NT;
  constexpr bool kCache = T <= 8;  // keep v, w, vn of the lane's columns in registers
  const TailJob jb = jobs[blockIdx.x];
  const int n = jb.m;
  double* ag = A + jb.a_off;
  double* d = Dg + jb.w_off;
  double* e = Eg + jb.w_off;
  double* tau = TAUg + jb.w_off;
  extern __shared__ double sm[];
  double* a = ag;
  int ld = jb.ld;
  double* sm2 = sm;
  if constexpr (kLds) {
    a = sm;
    ld = n;
    sm2 = sm + n * n;
    for (int q = threadIdx.x; q < n * n; q += NT) a[q] = ag[int64_t(q / n) * jb.ld + q % n];
    __syncthreads();
  }
  double* v = sm2;
  double* w = v + n;
  double* vn = w + n;
  double* prow = vn + n;
  double* pw = prow + n;      // NW x n column partials, summed in wave order
  double* red = pw + NW * n;  // 2 NW
  double* scal = red + 2 * NW;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int parity = 0;
  if (jb.pending < 0 && n <= 2) {
    if (tid == 0) {
      d[0] = a[0];
      e[0] = 0;
      tau[0] = 0;
      if (n == 2) {
        e[0] = a[1];
        d[1] = a[ld + 1];
        e[1] = 0;
        tau[1] = 0;
      }
    }
    return;
  }
  if (jb.pending < 0) {
    for (int j = tid; j < n; j += NT) {
      v[j] = 0;
      w[j] = 0;
    }
  } else {
    // the multi-workgroup kernel's last reflector (v[0] = 1) and p = A22 v:
    // w = tau p - tau^2/2 (p.v) v
    const double* vb = VP + int64_t(jb.pending) * w_n + jb.w_off;
    const double* pb = VP + int64_t(2 + jb.pending) * w_n + jb.w_off;
    const double tp = tau[-1];
    double loc = 0;
    for (int j = tid; j < n; j += NT) {
      const double vj = j == 0 ? 1.0 : vb[j];
      v[j] = vj;
      prow[j] = pb[j];
      loc += pb[j] * vj;
    }
    const double pv = block_sum<NT>(loc, red, parity);
    const double alw = -0.5 * tp * tp * pv;
    for (int j = tid; j < n; j += NT) w[j] = tp * prow[j] + alw * v[j];
  }
  __syncthreads();
  for (int k = 0; k + 2 < n; ++k) {
    // (a) row k with the pending update; sum of squares of x[1:]
    const double vk = v[k], wk = w[k];
    double* rowk = a + int64_t(k) * ld;
    double xr[R];
    double ss = 0;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int j = k + tid + u * NT;
      xr[u] = 0;
      if (j < n) {
        const double x = rowk[j] - (vk * w[j] + wk * v[j]);
        xr[u] = x;
        if (j == k) d[k] = x;
        if (j == k + 1) scal[0] = x;
        if (j >= k + 2) ss += x * x;
      }
    }
    const double sigma = block_sum<NT>(ss, red, parity);
    const double alpha = scal[0];
    double tk = 0, beta = alpha, scale = 0;
    if (sigma > 0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0 ? -nrm : nrm;
      tk = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int j = k + tid + u * NT;
      if (j < n && j > k) {
        const double vj = j == k + 1 ? 1.0 : xr[u] * scale;
        vn[j] = vj;
        if (j >= k + 2) rowk[j] = vj;
      }
    }
    if (tid == 0) {
      e[k] = beta;
      tau[k] = tk;
    }
    __syncthreads();
    // (c) rows k+1..n-1: pending update, then row and column parts of A22 vn
    double cacc[T];
    double vj[kCache ? T : 1], wj[kCache ? T : 1], vnj[kCache ? T : 1];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      cacc[t] = 0;
      if constexpr (kCache) {
        const int j = 64 * t + lane;
        const bool in = j > k && j < n;
        vj[t] = in ? v[j] : 0.0;
        wj[t] = in ? w[j] : 0.0;
        vnj[t] = in ? vn[j] : 0.0;
      }
    }
    const int first = k + 1;
    const int i0 = first + ((wv - first % NW) + NW) % NW;
    // two rows per wave and iteration, so twice the loads are in flight
    for (int i = i0; i < n; i += 2 * NW) {
      const int i2 = i + NW;
      const bool has2 = i2 < n;
      const double vi = v[i], wi = w[i], vni = vn[i];
      const double vi2 = has2 ? v[i2] : 0.0, wi2 = has2 ? w[i2] : 0.0, vni2 = has2 ? vn[i2] : 0.0;
      double* row = a + int64_t(i) * ld;
      double* row2 = a + int64_t(has2 ? i2 : i) * ld;
      double xv[T], xw[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int j = 64 * t + lane;
        xv[t] = (j >= i && j < n) ? row[j] : 0.0;
        xw[t] = (has2 && j >= i2 && j < n) ? row2[j] : 0.0;
      }
      double racc = 0, racc2 = 0;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int j = 64 * t + lane;
        if (j >= i && j < n) {
          double vjt, wjt, vnjt;
          if constexpr (kCache) {
            vjt = vj[t];
            wjt = wj[t];
            vnjt = vnj[t];
          } else {
            vjt = v[j];
            wjt = w[j];
            vnjt = vn[j];
          }
          const double x = xv[t] - (vi * wjt + wi * vjt);
          row[j] = x;
          racc += x * vnjt;
          if (j > i) cacc[t] += x * vni;
          if (has2 && j >= i2) {
            const double x2 = xw[t] - (vi2 * wjt + wi2 * vjt);
            row2[j] = x2;
            racc2 += x2 * vnjt;
            if (j > i2) cacc[t] += x2 * vni2;
          }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        racc += __shfl_xor(racc, o, 64);
        racc2 += __shfl_xor(racc2, o, 64);
      }
      if (lane == 0) {
        prow[i] = racc;
        if (has2) prow[i2] = racc2;
      }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int j = 64 * t + lane;
      if (j > k && j < n) pw[wv * n + j] = cacc[t];
    }
    __syncthreads();
    // (d) p = row part + column parts (fixed order); w = tau p - tau^2/2 (p.v) v; v <- vn
    double pv = 0;
    for (int j = k + 1 + tid; j < n; j += NT) {
      double p = prow[j];
#pragma unroll
      for (int q = 0; q < NW; ++q) p += pw[q * n + j];
      prow[j] = p;
      pv += p * vn[j];
    }
    pv = block_sum<NT>(pv, red, parity);
    const double alw = -0.5 * tk * tk * pv;
    for (int j = k + 1 + tid; j < n; j += NT) {
      w[j] = tk * prow[j] + alw * vn[j];
      v[j] = vn[j];
    }
    if (tid == 0) {
      v[k] = 0;
      w[k] = 0;
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int p = n - 2, q = n - 1;
    const double app = a[int64_t(p) * ld + p] - 2.0 * v[p] * w[p];
    const double apq = a[int64_t(p) * ld + q] - (v[p] * w[q] + w[p] * v[q]);
    const double aqq = a[int64_t(q) * ld + q] - 2.0 * v[q] * w[q];
    d[p] = app;
    e[p] = apq;
    d[q] = aqq;
    e[q] = 0;
    tau[p] = 0;
    tau[q] = 0;
  }
  if constexpr (kLds) {
    // reflector k is row k, columns k+2.. (the back-transformation's input)
    __syncthreads();
    for (int q = threadIdx.x; q < n * n; q += NT) {
      const int r = q / n, c = q % n;
      if (c >= r + 2) ag[int64_t(r) * jb.ld + c] = a[q];
    }
  }
}

// ---------------------------------------------------------------------------
// 1a'. The same trailing-block tridiagonalisation with the block in registers
// (m <= MT = 64 CG): wavefront wv holds rows wv + NW r (r < RPW) and lane l
// columns l + 64 g (g < CG) of the full symmetric block.  Per column step the
// owner of row k publishes it (pending update applied) to LDS together with
// the squared norm beyond the subdiagonal; every thread forms the reflector
// entries it needs from that row; each wavefront applies the pending rank-2 update to
// its rows and forms their parts of A22 v, reduced across the lanes through
// LDS (RPW partials per lane, summed by groups of lanes, no per-row
// reduction chains).  Reflector k goes to row k, columns k+2.., of A.
template <int NT, int MT>
__global__ __launch_bounds__(NT) void tridiag_reg_kernel(const TailJob* __restrict__ jobs, double* __restrict__ A,
                                                         double* __restrict__ Dg, double* __restrict__ Eg,
                                                         double* __restrict__ TAUg, const double* __restrict__ VP,
                                                         int64_t w_n) {
  constexpr int NW = NT / 64;
  constexpr int RPW = MT / NW;  // rows per wavefront
  constexpr int CG = MT / 64;   // column groups per lane
  // lanes summing one row's partials: the largest power of two <= 64 / RPW
  constexpr int LPR = 64 / RPW >= 8 ? 8 : 64 / RPW >= 4 ? 4 : 64 / RPW >= 2 ? 2 : 1;
  static_assert(MT % NW == 0 && MT % 64 == 0 && RPW <= 64, "tile shape");
  const TailJob jb = jobs[blockIdx.x];
  const int m = jb.m;
  double* ag = A + jb.a_off;
  const int ld = jb.ld;
  double* d = Dg + jb.w_off;
  double* e = Eg + jb.w_off;
  double* tau = TAUg + jb.w_off;
  extern __shared__ double sm[];
  double* v = sm;             // pending reflector (v[k] = 1 at its step)
  double* w = v + MT;         // pending w
  double* p = w + MT;         // A22 vn (vn: this step's reflector, formed from rowk)
  double* rowk = p + MT;      // row k with the pending update
  double* red = rowk + MT;    // 2 NW
  double* fin = red + 2 * NW; // the last 2 x 2 block
  double* part = fin + 8;     // NW x 64 x (RPW + 1) lane partials (padded against bank conflicts)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int parity = 0;
  if (jb.pending < 0 && m <= 2) {
    if (tid == 0) {
      d[0] = ag[0];
      e[0] = 0;
      tau[0] = 0;
      if (m == 2) {
        e[0] = ag[1];
        d[1] = ag[ld + 1];
        e[1] = 0;
        tau[1] = 0;
      }
    }
    return;
  }
  double x[RPW][CG];
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int i = wv + NW * r;
#pragma unroll
    for (int g = 0; g < CG; ++g) {
      const int j = lane + 64 * g;
      x[r][g] = (i < m && j < m) ? ag[int64_t(i) * ld + j] : 0.0;
    }
  }
  if (jb.pending < 0) {
    for (int j = tid; j < MT; j += NT) {
      v[j] = 0;
      w[j] = 0;
    }
  } else {
    // the multi-workgroup kernel's last reflector (v[0] = 1) and p = A22 v:
    // w = tau p - tau^2/2 (p.v) v
    const double* vb = VP + int64_t(jb.pending) * w_n + jb.w_off;
    const double* pb = VP + int64_t(2 + jb.pending) * w_n + jb.w_off;
    const double tp = tau[-1];
    double loc = 0;
    for (int j = tid; j < MT; j += NT) {
      const double vj = j < m ? (j == 0 ? 1.0 : vb[j]) : 0.0;
      const double pj = j < m ? pb[j] : 0.0;
      v[j] = vj;
      p[j] = pj;
      loc += pj * vj;
    }
    const double pv = block_sum<NT>(loc, red, parity);
    const double alw = -0.5 * tp * tp * pv;
    for (int j = tid; j < MT; j += NT) w[j] = tp * p[j] + alw * v[j];
  }
  __syncthreads();
  for (int k = 0; k + 2 < m; ++k) {
    // (a) the owner publishes row k with the pending update, and the squared
    // norm of its part beyond the subdiagonal (fin[4])
    if (wv == k % NW) {
      const int r = k / NW;
      const double vk = v[k], wk = w[k];
      double ss = 0;
#pragma unroll
      for (int rr = 0; rr < RPW; ++rr)
        if (rr == r) {
#pragma unroll
          for (int g = 0; g < CG; ++g) {
            const int j = lane + 64 * g;
            if (j >= k && j < m) {
              const double xk = x[rr][g] - (vk * w[j] + wk * v[j]);
              rowk[j] = xk;
              if (j >= k + 2) ss += xk * xk;
            }
          }
        }
      ss = wave_sum(ss);
      if (lane == 0) fin[4] = ss;
    }
    __syncthreads();
    // (b) reflector of row k; every thread forms the entries it needs from
    // rowk, so no barrier separates this from (c)
    const double sigma = fin[4];
    const double alpha = rowk[k + 1];
    double tk = 0, beta = alpha, scale = 0;
    if (sigma > 0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0 ? -nrm : nrm;
      tk = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    for (int j = k + 2 + tid; j < m; j += NT) ag[int64_t(k) * ld + j] = rowk[j] * scale;
    if (tid == 0) {
      d[k] = rowk[k];
      e[k] = beta;
      tau[k] = tk;
    }
    // this step's reflector entry j (j > k)
    auto refl = [&](int j) { return j == k + 1 ? 1.0 : rowk[j] * scale; };
    // (c) rows k+1..m-1: pending update, then the lane partials of A22 vn
    double vj[CG], wj[CG], vnj[CG];
#pragma unroll
    for (int g = 0; g < CG; ++g) {
      const int j = lane + 64 * g;
      const bool in = j > k && j < m;
      vj[g] = in ? v[j] : 0.0;
      wj[g] = in ? w[j] : 0.0;
      vnj[g] = in ? refl(j) : 0.0;
    }
    double* pw = part + (wv * 64 + lane) * (RPW + 1);
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int i = wv + NW * r;
      double acc = 0;
      if (i > k && i < m) {
        const double vi = v[i], wi = w[i];
#pragma unroll
        for (int g = 0; g < CG; ++g) {
          const double xx = x[r][g] - (vi * wj[g] + wi * vj[g]);
          x[r][g] = xx;
          acc += xx * vnj[g];
        }
      }
      pw[r] = acc;
    }
    __builtin_amdgcn_wave_barrier();
    {
      // lanes r * LPR .. r * LPR + LPR - 1 sum row r's 64 partials
      const int r = lane / LPR, q = lane % LPR;
      double s2 = 0;
      if (r < RPW) {
        const double* col = part + int64_t(wv * 64) * (RPW + 1) + r;
#pragma unroll
        for (int l = 0; l < 64 / LPR; ++l) s2 += col[(q * (64 / LPR) + l) * (RPW + 1)];
      }
#pragma unroll
      for (int o = 1; o < LPR; o <<= 1) s2 += __shfl_xor(s2, o, 64);
      const int i = wv + NW * r;
      const bool own = r < RPW && q == 0 && i > k && i < m;
      if (own) p[i] = s2;
      // this wavefront's rows' share of p.vn, summed over the wavefronts below
      const double pvw = wave_sum(own ? s2 * refl(i) : 0.0);
      if (lane == 0) red[parity * NW + wv] = pvw;
    }
    __syncthreads();
    // (d) w = tau p - tau^2/2 (p.vn) vn; v <- vn
    double pv = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) pv += red[parity * NW + q];
    parity ^= 1;
    const double alw = -0.5 * tk * tk * pv;
    for (int j = k + 1 + tid; j < m; j += NT) {
      const double vnj2 = refl(j);
      w[j] = tk * p[j] + alw * vnj2;
      v[j] = vnj2;
    }
    if (tid == 0) {
      v[k] = 0;
      w[k] = 0;
    }
    __syncthreads();
  }
  // the last 2 x 2 block with the pending update
  const int pq[2] = {m - 2, m - 1};
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int i = pq[a];
    if (wv == i % NW) {
      const int r = i / NW;
#pragma unroll
      for (int rr = 0; rr < RPW; ++rr)
        if (rr == r) {
#pragma unroll
          for (int g = 0; g < CG; ++g) {
            const int j = lane + 64 * g;
            if (j == m - 2 || j == m - 1) fin[a * 2 + (j - (m - 2))] = x[rr][g];
          }
        }
    }
  }
  __syncthreads();
  if (tid == 0) {
    const int p0 = m - 2, q0 = m - 1;
    d[p0] = fin[0] - 2.0 * v[p0] * w[p0];
    e[p0] = fin[1] - (v[p0] * w[q0] + w[p0] * v[q0]);
    d[q0] = fin[3] - 2.0 * v[q0] * w[q0];
    e[q0] = 0;
    tau[p0] = 0;
    tau[q0] = 0;
  }
}

// ---------------------------------------------------------------------------
// 1b. Tridiagonalisation for large matrices: one launch per column step k,
// batched over the jobs, each workgroup of R wavefronts owning R * RW rows of
// one job (RW rows per wavefront) of the full symmetric matrix.  Every
// workgroup recomputes the O(n) quantities it needs: the pending rank-2
// update of step k-1 (from the previous step's A22 v, double-buffered) and
// the reflector of step k (from row k, which no workgroup writes during the
// step).  Reflector k is stored in column k below the subdiagonal by the
// owners of those rows and moved into row k afterwards.  VP holds per job: v
// (two buffers) and A22 v (two buffers), each indexed like the eigenvalues.
// Workgroups of a job at column step k (R here = rows per workgroup); the
// job's column steps stop at k0 (DevJob::pad), where the single-workgroup
// kernel takes over its tail.
template <int R>
__device__ __forceinline__ int step_count(int n, int k, int k0) {
  return k < k0 ? (n - k - 1 + R - 1) / R : 0;
}

// MaxN bounds the trailing size n - k of every job in the launch (the
// register arrays below hold the wave's rows and the step's vectors); each
// wavefront owns RW rows (i, i + R, ...), so a workgroup covers R * RW rows
// and shares its per-step preamble (job search, row k, reflector) among them.
template <int R, int MaxN, int RW>
__global__ __launch_bounds__(64 * R) void tdstep_kernel(const DevJob* __restrict__ jobs, int n_jobs, int k,
                                                        double* __restrict__ A, double* __restrict__ Dg,
                                                        double* __restrict__ Eg, double* __restrict__ TAUg,
                                                        double* __restrict__ VP, int64_t w_n, int ldsn) {
  constexpr int kStepThreads = 64 * R;
  constexpr int kStepRows = R * RW;
  constexpr int kPerThread = (MaxN + kStepThreads - 1) / kStepThreads;  // row-k / vector entries per thread
  constexpr int kPerLane = MaxN / 64;                                     // own-row entries per lane
  extern __shared__ double sm[];
  double* vprev = sm;
  double* wprev = sm + ldsn;
  double* vcur = sm + 2 * ldsn;
  __shared__ double red[2 * R];
  // this workgroup's job: every wavefront scans the job list 64 entries at a
  // time (one load per lane, an inclusive prefix sum of the per-job
  // workgroup counts, a ballot), instead of one dependent load per job
  int id = int(blockIdx.x), q = n_jobs;
  {
    const int ln = int(threadIdx.x & 63);
    int base = 0;
    for (int q0 = 0; q0 < n_jobs; q0 += 64) {
      const int qq = q0 + ln;
      const int cnt = qq < n_jobs ? step_count<kStepRows>(jobs[qq].n, k, jobs[qq].pad) : 0;
      int incl = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (ln >= o) incl += t;
      }
      const int total = __shfl(incl, 63, 64);
      if (id < base + total) {
        const unsigned long long hit = __ballot(base + incl > id);
        const int l = __ffsll(static_cast<long long>(hit)) - 1;
        q = q0 + l;
        id -= base + __shfl(incl - cnt, l, 64);
        break;
      }
      base += total;
    }
  }
  if (q >= n_jobs) return;
  const DevJob jb = jobs[q];
  const int n = jb.n, chunk = id;
  double* a = A + jb.a_off;
  double* d = Dg + jb.w_off;
  double* e = Eg + jb.w_off;
  double* tau = TAUg + jb.w_off;
  const double* vb_prev = VP + int64_t((k + 1) & 1) * w_n + jb.w_off;
  double* vb_cur = VP + int64_t(k & 1) * w_n + jb.w_off;
  const double* pb_prev = VP + int64_t(2 + ((k + 1) & 1)) * w_n + jb.w_off;
  double* pb_cur = VP + int64_t(2 + (k & 1)) * w_n + jb.w_off;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int parity = 0;
  if (n == 1) {
    if (tid == 0) {
      d[0] = a[0];
      e[0] = 0;
      tau[0] = 0;
    }
    return;
  }
  const bool pend = k >= 1;
  const bool last = k + 2 == n;
  const int i0 = k + 1 + chunk * kStepRows + wv;  // this wavefront's rows: i0 + R * rr
  // every global operand of the step is requested up front
  double pp[kPerThread], vv[kPerThread], xk[kPerThread], xr[RW][kPerLane];
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int j = k + tid + u * kStepThreads;
    const bool in = j < n;
    pp[u] = (pend && in) ? pb_prev[j] : 0.0;
    vv[u] = (pend && in) ? (j == k ? 1.0 : vb_prev[j]) : 0.0;
    xk[u] = (!last && in) ? a[int64_t(k) * n + j] : 0.0;
  }
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int i = i0 + R * rr;
    const bool own = !last && i < n;
    const double* rowi = a + int64_t(own ? i : 0) * n;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
      const int j = k + 1 + lane + 64 * u;
      xr[rr][u] = (own && j < n) ? rowi[j] : 0.0;
    }
  }
  // pending update of step k-1: w = tau p - tau^2/2 (p.v) v over j >= k
  if (pend) {
    const double tp = tau[k - 1];
    double loc = 0;
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int j = k + tid + u * kStepThreads;
      if (j < n) {
        vprev[j] = vv[u];
        loc += pp[u] * vv[u];
      }
    }
    const double pv = block_sum<kStepThreads>(loc, red, parity);
    const double alw = -0.5 * tp * tp * pv;
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int j = k + tid + u * kStepThreads;
      if (j < n) wprev[j] = tp * pp[u] + alw * vv[u];
    }
    __syncthreads();
  }
  if (last) {  // the final 2 x 2 block
    if (tid == 0) {
      const int p0 = k, p1 = k + 1;
      double app = a[int64_t(p0) * n + p0], apq = a[int64_t(p0) * n + p1], aqq = a[int64_t(p1) * n + p1];
      if (pend) {
        app -= 2.0 * vprev[p0] * wprev[p0];
        apq -= vprev[p0] * wprev[p1] + wprev[p0] * vprev[p1];
        aqq -= 2.0 * vprev[p1] * wprev[p1];
      }
      d[p0] = app;
      e[p0] = apq;
      d[p1] = aqq;
      e[p1] = 0;
      tau[p0] = 0;
      tau[p1] = 0;
    }
    return;
  }
  // row k with the pending update, then its reflector
  const double vk = pend ? vprev[k] : 0.0, wk = pend ? wprev[k] : 0.0;
  double ss = 0;
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int j = k + tid + u * kStepThreads;
    if (j < n) {
      double x = xk[u];
      if (pend) x -= vk * wprev[j] + wk * vprev[j];
      xk[u] = x;
      vcur[j] = x;
      if (j >= k + 2) ss += x * x;
    }
  }
  const double sigma = block_sum<kStepThreads>(ss, red, parity);
  const double alpha = vcur[k + 1], dk = vcur[k];
  double tk = 0, beta = alpha, scale = 0;
  if (sigma > 0) {
    const double nrm = sqrt(alpha * alpha + sigma);
    beta = alpha >= 0 ? -nrm : nrm;
    tk = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int j = k + tid + u * kStepThreads;
    if (j > k && j < n) vcur[j] = j == k + 1 ? 1.0 : xk[u] * scale;
  }
  __syncthreads();
  if (chunk == 0 && tid == 0) {
    d[k] = dk;
    e[k] = beta;
    tau[k] = tk;
  }
  // this wavefront's rows: pending update over columns >= k+1, then (A22 v)_i
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int i = i0 + R * rr;
    if (i >= n) break;
    double* row = a + int64_t(i) * n;
    const double vi = pend ? vprev[i] : 0.0, wi = pend ? wprev[i] : 0.0;
    double acc = 0;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
      const int j = k + 1 + lane + 64 * u;
      if (j < n) {
        double x = xr[rr][u];
        if (pend) x -= vi * wprev[j] + wi * vprev[j];
        row[j] = x;
        acc += x * vcur[j];
      }
    }
    acc = wave_sum(acc);
    if (lane == 0) {
      pb_cur[i] = acc;
      vb_cur[i] = vcur[i];
      if (i >= k + 2) row[k] = vcur[i];
    }
  }
}

template <int R, int MaxN, int RW>
void launch_tdstep_t(dim3 grid, size_t lds, hipStream_t st, const DevJob* jobs, int nj, int k, double* A, double* D,
                     double* E, double* TAU, double* VP, int64_t w_n, int ldsn) {
  hipLaunchKernelGGL((tdstep_kernel<R, MaxN, RW>), grid, dim3(64 * R), lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n,
                     ldsn);
}

template <int R, int RW>
void launch_tdstep_r(int cap, dim3 grid, size_t lds, hipStream_t st, const DevJob* jobs, int nj, int k, double* A,
                     double* D, double* E, double* TAU, double* VP, int64_t w_n, int ldsn) {
  if (cap <= 512)
    launch_tdstep_t<R, 512, RW>(grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
  else if (cap <= 1024)
    launch_tdstep_t<R, 1024, RW>(grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
  else
    launch_tdstep_t<R, kSymEigMaxN, RW>(grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
}

// one column step over every active job: R waves per workgroup, rw rows per
// wave, registers sized for trailing blocks of at most `cap` (512, 1024 or 2048)
void launch_tdstep(int r, int rw, int cap, dim3 grid, size_t lds, hipStream_t st, const DevJob* jobs, int nj, int k,
                   double* A, double* D, double* E, double* TAU, double* VP, int64_t w_n, int ldsn) {
  if (rw == 2)
    launch_tdstep_r<8, 2>(cap, grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
  else if (r == 16)
    launch_tdstep_r<16, 1>(cap, grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
  else if (r == 4)
    launch_tdstep_r<4, 1>(cap, grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
  else
    launch_tdstep_r<8, 1>(cap, grid, lds, st, jobs, nj, k, A, D, E, TAU, VP, w_n, ldsn);
}

// Reflector k < k0 (DevJob::pad) from column k (rows >= k+2) into row k
// (columns >= k+2), for the back-transformation; 32 x 32 tiles of the
// strictly lower part.
struct TileItem {
  int32_t job, ti, tj, pad;
};
__global__ __launch_bounds__(256) void reflector_rows_kernel(const TileItem* __restrict__ items,
                                                             const DevJob* __restrict__ jobs, double* __restrict__ A) {
  __shared__ double t[32][33];
  const TileItem it = items[blockIdx.x];
  const DevJob jb = jobs[it.job];
  const int n = jb.n;
  double* a = A + jb.a_off;
  const int r0 = it.ti * 32, c0 = it.tj * 32;  // source tile rows r0.., columns c0.. (r0 >= c0)
  for (int e2 = threadIdx.x; e2 < 32 * 32; e2 += 256) {
    const int r = e2 >> 5, c = e2 & 31;
    const int gi = r0 + r, gj = c0 + c;
    t[r][c] = (gi < n && gj < n) ? a[int64_t(gi) * n + gj] : 0.0;
  }
  __syncthreads();
  for (int e2 = threadIdx.x; e2 < 32 * 32; e2 += 256) {
    const int r = e2 >> 5, c = e2 & 31;  // destination row c0 + r, column r0 + c
    const int gk = c0 + r, gj = r0 + c;
    if (gk < jb.pad && gj < n && gj >= gk + 2 && gk + 2 < n) a[int64_t(gk) * n + gj] = t[c][r];
  }
}

// ---------------------------------------------------------------------------
// 2. Bisection: thread j of item (job, j0) finds eigenvalue j0 + j (ascending
// index) to max(2 eps |lambda|, 2 eps |T|) absolute width.
struct BisItem {
  int32_t job, j0;
};

template <int NT, int kBisLanes>
__global__ __launch_bounds__(NT) void bisect_kernel(const BisItem* __restrict__ items, const DevJob* __restrict__ jobs,
                                                    const double* __restrict__ Dg, const double* __restrict__ Eg,
                                                    double* __restrict__ W, double* __restrict__ TN, int descending) {
  constexpr int NW = NT / 64;
  const BisItem it = items[blockIdx.x];
  const DevJob jb = jobs[it.job];
  const int n = jb.n;
  extern __shared__ double sm[];
  double* d = sm;
  double* e2 = d + n;
  __shared__ double rmin[NW], rmax[NW], re2[NW];
  const double* dg = Dg + jb.w_off;
  const double* eg = Eg + jb.w_off;
  const int tid = threadIdx.x;
  double lo_l = INFINITY, hi_l = -INFINITY, emax = 0;
  for (int i = tid; i < n; i += NT) {
    const double di = dg[i];
    const double el = i > 0 ? fabs(eg[i - 1]) : 0.0;
    const double er = i + 1 < n ? fabs(eg[i]) : 0.0;
    d[i] = di;
    if (i + 1 < n) {
      e2[i] = eg[i] * eg[i];
      emax = fmax(emax, e2[i]);
    }
    lo_l = fmin(lo_l, di - el - er);
    hi_l = fmax(hi_l, di + el + er);
  }
  lo_l = wave_min(lo_l);
  hi_l = wave_max(hi_l);
  emax = wave_max(emax);
  if ((tid & 63) == 0) {
    rmin[tid >> 6] = lo_l;
    rmax[tid >> 6] = hi_l;
    re2[tid >> 6] = emax;
  }
  __syncthreads();
  double gl = rmin[0], gu = rmax[0], e2max = re2[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) {
    gl = fmin(gl, rmin[i]);
    gu = fmax(gu, rmax[i]);
    e2max = fmax(e2max, re2[i]);
  }
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = DBL_MIN * fmax(1.0, e2max);
  gl -= 4.0 * kEps * tnorm + 2.0 * pivmin;
  gu += 4.0 * kEps * tnorm + 2.0 * pivmin;
  if (it.j0 == 0 && tid == 0) TN[it.job] = tnorm;
  // multisection: kBisLanes lanes per eigenvalue evaluate kBisLanes interior
  // points of the bracket per round and keep the sub-interval of the target
  const int g = (tid & 63) / kBisLanes, t = tid & (kBisLanes - 1);
  const int j = it.j0 + (tid / kBisLanes);
  const bool valid = j < n;
  const double atol = 2.0 * kEps * tnorm;
  double lo = gl, hi = gu;
  for (int round = 0; round < 64; ++round) {
    const bool done = !valid || hi - lo <= fmax(atol, 2.0 * kEps * fmax(fabs(lo), fabs(hi)));
    const double x = lo + (hi - lo) * double(t + 1) / double(kBisLanes + 1);
    const bool inside = !done && x > lo && x < hi;
    const int cnt = inside ? sturm_count(d, e2, n, x, pivmin) : 0;
    const unsigned long long above = __ballot(inside && cnt > j);
    const unsigned long long any_in = __ballot(inside);
    const unsigned gm = unsigned((above >> (g * kBisLanes)) & ((1ull << kBisLanes) - 1));
    const unsigned gi = unsigned((any_in >> (g * kBisLanes)) & ((1ull << kBisLanes) - 1));
    double nlo = lo, nhi = hi;
    if (gm != 0u) {
      const int f = __ffs(int(gm)) - 1;
      nhi = __shfl(x, g * kBisLanes + f, 64);
      if (f > 0) nlo = __shfl(x, g * kBisLanes + f - 1, 64);
    } else {
      const int top = 31 - __clz(int(gi | 1u));
      const double xt = __shfl(x, g * kBisLanes + top, 64);
      if (gi != 0u) nlo = xt;
    }
    if (gi != 0u) {
      lo = nlo;
      hi = nhi;
    }
    if (__ballot(gi != 0u) == 0ull) break;
  }
  if (valid && t == 0) W[jb.w_off + (descending ? n - 1 - j : j)] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------------------
// 3a. Inverse iteration, one thread per wanted vector: T - lambda I = P L U
// with partial pivoting (U has two superdiagonals), two solves from a
// hashed start, each normalised by its largest entry; finally unit 2-norm.
//
// Equal eigenvalues: an eigenvalue of multiplicity m of the reduced matrix
// makes T split into unreduced blocks (off-diagonals at the rounding level,
// |e_i| <= kSplitTol eps |T|), each holding one copy.  Inverse iteration
// with the common shift on the whole of T would send every copy to the
// block whose eigenvalue happens to lie closest, and orthogonalising those
// m nearly parallel vectors leaves rounding noise.  So for a run of
// eigenvalues within kClusterTol eps |T| of each other, the off-diagonals at
// the split points are taken as zero, the blocks holding the run's values
// are counted by Sturm sequences, and the run's r-th vector starts (and, T
// being block diagonal then, stays) on the r-th of them.
constexpr double kSplitTol = 16.0;
constexpr double kClusterTol = 64.0;

struct VecItem {
  int32_t job, c;
};

// Eigenvalues below x of the unreduced block rows [a, b) of T with the
// split off-diagonals (|e_i| <= tol_e) taken as zero.
__device__ int sturm_count_block(const double* d, const double* e, int a, int b, double x, double pivmin) {
  double q = d[a] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int cnt = q < 0.0;
  for (int i = a + 1; i < b; ++i) {
    q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
    if (fabs(q) <= pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Factors of T - lambda I for one vector, kLuFields = 4 doubles per (row i,
// vector c) at lu[(i * k + c) * 4]: U0 (pivot), U1, U2 (superdiagonals) and
// the multiplier.  Row i was interchanged with row i + 1 when U2 is nonzero
// or -0.0; an unswapped row stores U2 = +0.0.
__device__ __forceinline__ bool lu_swapped(double u2) { return u2 != 0.0 || signbit(u2); }

// One solve with the factors of vector c stored by invit_kernel: y <- U^-1
// L^-1 P y for the vector y[i * ys], in place, by the calling thread; returns
// max |y|.
__device__ double lu_solve_inplace(const double* __restrict__ lu, int k, int c, int n, double* y, int ys) {
  auto F = [&](int i) { return lu + (int64_t(i) * k + c) * kLuFields; };
  double cur = y[0];
  for (int i = 0; i + 1 < n; ++i) {
    double nxt = y[(i + 1) * ys];
    if (lu_swapped(F(i)[2])) {
      const double t = cur;
      cur = nxt;
      nxt = t;
    }
    nxt -= F(i)[3] * cur;
    y[i * ys] = cur;
    cur = nxt;
  }
  y[(n - 1) * ys] = cur;
  double z1 = 0, z2 = 0, zmax = 0;
  for (int i = n - 1; i >= 0; --i) {
    const double zi = (y[i * ys] - F(i)[1] * z1 - F(i)[2] * z2) / F(i)[0];
    y[i * ys] = zi;
    zmax = fmax(zmax, fabs(zi));
    z2 = z1;
    z1 = zi;
  }
  return zmax;
}

// The solves are serial in the rows, one thread per vector, and few vectors
// run at once, so the loads' latency sets the time: every pass loads rows
// in chunks of C, the factors of a row as two 16-byte loads, and there are
// five passes over the rows: the factorisation runs the first forward solve
// on the hashed start as it goes, and the last back substitution sums the
// vector's 2-norm as it goes.
__global__ __launch_bounds__(64) void invit_kernel(const VecItem* __restrict__ items, int n_items,
                                                   const DevJob* __restrict__ jobs, const double* __restrict__ Dg,
                                                   const double* __restrict__ Eg, const double* __restrict__ W,
                                                   const double* __restrict__ TN, double* __restrict__ LU,
                                                   double* __restrict__ Z) {
  constexpr int C = 8;  // rows per prefetched chunk
  const int g = blockIdx.x * 64 + threadIdx.x;
  if (g >= n_items) return;
  const VecItem it = items[g];
  const DevJob jb = jobs[it.job];
  const int n = jb.n, k = jb.k, c = it.c;
  const double* d = Dg + jb.w_off;
  const double* e = Eg + jb.w_off;
  const double lam = W[jb.w_off + c];
  double* lu = LU + jb.lu_off;
  double* z = Z + jb.z_off;
  auto F = [&](int i) { return reinterpret_cast<double2*>(lu + (int64_t(i) * k + c) * kLuFields); };
  auto Y = [&](int i) -> double& { return z[int64_t(i) * k + c]; };
  if (n == 1) {
    Y(0) = 1.0;
    return;
  }
  const double tn = fmax(TN[jb.job], DBL_MIN);
  const double piv = kEps * tn;
  auto guard = [&](double x) { return fabs(x) < piv ? (x < 0 ? -piv : piv) : x; };
  // a run of equal eigenvalues: this vector's split block [s0, s1)
  const double tol_e = kSplitTol * kEps * tn;
  bool split = false;
  int s0 = 0, s1 = n;
  {
    const double* w = W + jb.w_off;
    const double tol_c = kClusterTol * kEps * tn;
    int r0 = c, r1 = c;
    while (r0 > 0 && fabs(w[r0 - 1] - w[r0]) <= tol_c) --r0;
    while (r1 + 1 < n && fabs(w[r1 + 1] - w[r1]) <= tol_c) ++r1;
    if (r1 > r0) {
      const double lo = fmin(w[r0], w[r1]) - tol_c, hi = fmax(w[r0], w[r1]) + tol_c;
      double e2max = 0;
      for (int i = 0; i + 1 < n; ++i) e2max = fmax(e2max, e[i] * e[i]);
      const double pivmin = DBL_MIN * fmax(1.0, e2max);
      const int rank = c - r0;
      int a = 0, cum = 0, p0 = -1, p1 = -1;
      for (int i = 0; i < n; ++i) {
        if (i + 1 < n && fabs(e[i]) > tol_e) continue;
        const int cnt = sturm_count_block(d, e, a, i + 1, hi, pivmin) - sturm_count_block(d, e, a, i + 1, lo, pivmin);
        if (p0 < 0 && rank < cum + cnt) {
          p0 = a;
          p1 = i + 1;
        }
        cum += cnt;
        a = i + 1;
      }
      if (cum == r1 - r0 + 1 && p0 >= 0) {
        split = true;
        s0 = p0;
        s1 = p1;
      }
    }
  }
  auto Ev = [&](int i) { return split && fabs(e[i]) <= tol_e ? 0.0 : e[i]; };
  const uint32_t seed = uint32_t(c) * 2654435761u + uint32_t(jb.job);
  auto start = [&](int i) { return i >= s0 && i < s1 ? hash_unit(seed, uint32_t(i)) : 0.0; };
  // T - lambda I = P L U with partial pivoting, and L^-1 P applied to the
  // hashed start on the way (cur: the running entry of the forward solve)
  double c0 = d[0] - lam, c1 = Ev(0);
  double cur = start(0);
  for (int i0 = 0; i0 + 1 < n; i0 += C) {
    double dn[C], ei[C], en[C];
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const int i = i0 + u;
      dn[u] = i + 1 < n ? d[i + 1] : 0.0;
      ei[u] = i + 1 < n ? Ev(i) : 0.0;
      en[u] = i + 2 < n ? Ev(i + 1) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < C; ++u) {
      const int i = i0 + u;
      if (i + 1 >= n) break;
      const double dd = dn[u] - lam;
      const double yn = start(i + 1);
      if (fabs(c0) >= fabs(ei[u])) {
        c0 = guard(c0);
        const double m = fast_quot(ei[u], c0);
        F(i)[0] = make_double2(c0, c1);
        F(i)[1] = make_double2(0.0, m);
        c0 = dd - m * c1;
        c1 = en[u];
        Y(i) = cur;
        cur = yn - m * cur;
      } else {
        // |c0| < |e_i|; below the normal range the hardware reciprocal of e_i
        // overflows, so that rare case takes the IEEE quotient (|m| < 1)
        const double m = fabs(ei[u]) >= DBL_MIN ? fast_quot(c0, ei[u]) : c0 / ei[u];
        F(i)[0] = make_double2(ei[u], dd);
        F(i)[1] = make_double2(en[u] == 0.0 ? -0.0 : en[u], m);
        c0 = c1 - m * dd;
        c1 = -m * en[u];
        Y(i) = yn;
        cur = cur - m * yn;
      }
    }
  }
  F(n - 1)[0] = make_double2(guard(c0), 0.0);
  F(n - 1)[1] = make_double2(0.0, 0.0);
  Y(n - 1) = cur;
  double s = 1.0, sumsq = 0.0;
  for (int itn = 0; itn < 2; ++itn) {
    if (itn > 0) {
      // forward: row interchanges and multipliers applied to s * y
      double cf = s * Y(0);
      for (int i0 = 0; i0 + 1 < n; i0 += C) {
        double yn[C], u2[C], ml[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
          const int i = i0 + u;
          const bool in = i + 1 < n;
          const double2 b = in ? F(i)[1] : make_double2(0.0, 0.0);
          yn[u] = in ? Y(i + 1) : 0.0;
          u2[u] = b.x;
          ml[u] = b.y;
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
          const int i = i0 + u;
          if (i + 1 >= n) break;
          double nxt = s * yn[u];
          if (lu_swapped(u2[u])) {
            const double t = cf;
            cf = nxt;
            nxt = t;
          }
          nxt -= ml[u] * cf;
          Y(i) = cf;
          cf = nxt;
        }
      }
      Y(n - 1) = cf;
    }
    // back substitution with the two superdiagonals
    double z1 = 0, z2 = 0, zmax = 0, ss = 0;
    for (int i1 = n - 1; i1 >= 0; i1 -= C) {
      double yv[C], u0[C], u1[C], u2[C];
#pragma unroll
      for (int u = 0; u < C; ++u) {
        const int i = i1 - u;
        const bool in = i >= 0;
        const double2 a = in ? F(i)[0] : make_double2(1.0, 0.0);
        const double2 b = in ? F(i)[1] : make_double2(0.0, 0.0);
        yv[u] = in ? Y(i) : 0.0;
        u0[u] = a.x;
        u1[u] = a.y;
        u2[u] = b.x;
      }
#pragma unroll
      for (int u = 0; u < C; ++u) {
        const int i = i1 - u;
        if (i < 0) break;
        const double zi = fast_quot(yv[u] - u1[u] * z1 - u2[u] * z2, u0[u]);
        Y(i) = zi;
        zmax = fmax(zmax, fabs(zi));
        ss += zi * zi;
        z2 = z1;
        z1 = zi;
      }
    }
    s = zmax > 0 ? 1.0 / zmax : 1.0;
    sumsq = ss;
  }
  // unit 2-norm; the sum of squares of the unscaled solve is exact enough
  // unless it left the finite range, which a scaled pass then recomputes
  double f;
  if (sumsq <= 1e300) {
    f = sumsq > 0 ? 1.0 / sqrt(sumsq) : 1.0;
  } else {
    double nrm = 0;
    for (int i = 0; i < n; ++i) nrm += (Y(i) * s) * (Y(i) * s);
    f = nrm > 0 ? s / sqrt(nrm) : 1.0;
  }
  constexpr int C2 = 16;
  for (int i0 = 0; i0 < n; i0 += C2) {
    double yv[C2];
#pragma unroll
    for (int u = 0; u < C2; ++u) yv[u] = i0 + u < n ? Y(i0 + u) : 0.0;
#pragma unroll
    for (int u = 0; u < C2; ++u)
      if (i0 + u < n) Y(i0 + u) = yv[u] * f;
  }
}

// Column 2-norms of the n x k row-major Z of every job into norms (at w_off).
__global__ __launch_bounds__(256) void colnorm_kernel(const DevJob* __restrict__ jobs, const double* __restrict__ Z,
                                                      double* __restrict__ norms) {
  // 16 columns per workgroup, the rows split over 16 lanes per column (a
  // row's 16 columns are one 128-byte segment); the 16 partial sums of a
  // column are added in lane order
  __shared__ double part[16][17];
  const DevJob jb = jobs[blockIdx.y];
  const int cl = threadIdx.x & 15, rl = threadIdx.x >> 4;
  const int c = blockIdx.x * 16 + cl;
  const double* z = Z + jb.z_off;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (c < jb.k) {
    int i = rl;
    for (; i + 48 < jb.n; i += 64) {
      const double x0 = z[int64_t(i) * jb.k + c], x1 = z[int64_t(i + 16) * jb.k + c];
      const double x2 = z[int64_t(i + 32) * jb.k + c], x3 = z[int64_t(i + 48) * jb.k + c];
      s0 += x0 * x0;
      s1 += x1 * x1;
      s2 += x2 * x2;
      s3 += x3 * x3;
    }
    for (; i < jb.n; i += 16) {
      const double x = z[int64_t(i) * jb.k + c];
      s0 += x * x;
    }
  }
  part[rl][cl] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (threadIdx.x < 16 && c < jb.k) {
    double t = 0;
    for (int r = 0; r < 16; ++r) t += part[r][cl];
    norms[jb.w_off + c] = sqrt(t);
  }
}

// ---------------------------------------------------------------------------
// 3b. Orthonormalisation of the columns [c_begin, c_begin + c_count) of the
// n x k row-major Z (ld k) of each job (c_begin = the job's first column,
// DevJob::pad, + step * c_count), in column order, the columns before
// c_begin being orthonormal and already projected out: each block of NB
// columns is projected against the earlier columns of the range (two passes,
// 64 columns at a time), then orthonormalised within the block (two passes
// per column).  A column that keeps less than 1e-3 of its norm on entry
// (norms, at w_off of the job) was numerically dependent on the earlier ones;
// it restarts from a hashed vector projected against every earlier column.
template <int NT, int NB>
__global__ __launch_bounds__(NT) void orth_kernel(const DevJob* __restrict__ jobs, double* __restrict__ Z,
                                                  const double* __restrict__ LU, int step, int c_count,
                                                  const double* __restrict__ norms, const int* __restrict__ done) {
  constexpr int NW = NT / 64;
  if (done != nullptr && done[blockIdx.x] != 0) return;  // block finished by Cholesky QR
  const DevJob jb = jobs[blockIdx.x];
  const int n = jb.n, k = jb.k;
  const int c_begin = jb.pad + step * c_count;  // DevJob::pad: the first column to process
  if (k <= c_begin) return;
  const int c_end = min(k, c_begin + c_count);
  double* z = Z + jb.z_off;
  extern __shared__ double sm[];
  double* Zb = sm;                      // n x NB: the current block
  double* Hp = Zb + int64_t(n) * NB;    // NW x 64 x NB partial projections
  double* H = Hp + NW * 64 * NB;        // 64 x NB projections
  double* red = H + 64 * NB;            // NW x NB
  double* hs = red + NW * NB;           // NB workgroup sums
  double* nrm0 = hs + NB;               // NB norms on entry
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  // hs[c] = sum over the workgroup of vals[c], c < nv
  auto block_sums = [&](double(&vals)[NB], int nv) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < nv) {
        const double s = wave_sum(vals[c]);
        if (lane == 0) red[wv * NB + c] = s;
      }
    __syncthreads();
    if (tid < nv) {
      double s = 0;
      for (int q = 0; q < NW; ++q) s += red[q * NB + tid];
      hs[tid] = s;
    }
    __syncthreads();
  };
  // one classical pass: columns [c_lo, c_hi) of Zb minus their projection on
  // z[:, p_lo:p_end), 64 earlier columns at a time
  auto project = [&](int p_lo, int p_end, int c_lo, int c_hi) {
    for (int p0 = p_lo; p0 < p_end; p0 += 64) {
      const int pc = min(64, p_end - p0);
      double acc[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) acc[c] = 0;
      if (lane < pc) {
        // four rows in flight per wavefront
        const double* zc = z + p0 + lane;
        int i = wv;
        for (; i + 3 * NW < n; i += 4 * NW) {
          double zq[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) zq[q] = zc[int64_t(i + q * NW) * k];
#pragma unroll
          for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < NB; ++c)
              if (c >= c_lo && c < c_hi) acc[c] += zq[q] * Zb[(i + q * NW) * NB + c];
        }
        for (; i < n; i += NW) {
          const double zp = zc[int64_t(i) * k];
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c >= c_lo && c < c_hi) acc[c] += zp * Zb[i * NB + c];
        }
      }
#pragma unroll
      for (int c = 0; c < NB; ++c) Hp[(wv * 64 + lane) * NB + c] = acc[c];
      __syncthreads();
      for (int e = tid; e < 64 * NB; e += NT) {
        double s = 0;
        for (int q = 0; q < NW; ++q) s += Hp[q * 64 * NB + e];
        H[e] = s;
      }
      __syncthreads();
      for (int i = tid; i < n; i += NT) {
        double upd[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) upd[c] = 0;
        // eight earlier columns of the row in flight
        const double* zr = z + int64_t(i) * k + p0;
        int p = 0;
        for (; p + 8 <= pc; p += 8) {
          double zq[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) zq[q] = zr[p + q];
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int c = 0; c < NB; ++c) upd[c] += zq[q] * H[(p + q) * NB + c];
        }
        for (; p < pc; ++p) {
          const double zp = zr[p];
#pragma unroll
          for (int c = 0; c < NB; ++c) upd[c] += zp * H[p * NB + c];
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c >= c_lo && c < c_hi) Zb[i * NB + c] -= upd[c];
      }
      __syncthreads();
    }
  };

  // Cholesky QR of the nb columns in Zb: G = Zb^T Zb, G = R^T R, Zb <- Zb R^-1.
  // With `test`, fails (Zb untouched) when some R_cc <= 1e-3 nrm0[c].
  auto cholqr = [&](int nb, bool test) -> bool {
    constexpr int NU = NB * (NB + 1) / 2;
    double gp[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) gp[u] = 0;
    for (int i = tid; i < n; i += NT) {
      double zr[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) zr[c] = Zb[i * NB + c];
      int u = 0;
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = a; b < NB; ++b) gp[u++] += zr[a] * zr[b];
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const double x = wave_sum(gp[u]);
      if (lane == 0) Hp[wv * NU + u] = x;
    }
    __syncthreads();
    if (tid < NU) {
      double x = 0;
      for (int q = 0; q < NW; ++q) x += Hp[q * NU + tid];
      H[tid] = x;
    }
    __syncthreads();
    double* R = H + 64;  // NB x NB, upper triangle
    if (tid == 0) {
      bool ok = true;
      double minr = 1.0;
      double g[NB][NB];
      int u = 0;
      for (int a = 0; a < NB; ++a)
        for (int b = a; b < NB; ++b) {
          g[a][b] = H[u];
          g[b][a] = H[u];
          ++u;
        }
      for (int a = 0; a < nb && ok; ++a) {
        double d = g[a][a];
        for (int c = 0; c < a; ++c) d -= R[c * NB + a] * R[c * NB + a];
        if (!(d > 0) || (test && sqrt(d) <= 1e-3 * nrm0[a])) {
          ok = false;
          break;
        }
        const double r = sqrt(d);
        if (test) minr = fmin(minr, r / fmax(nrm0[a], DBL_MIN));
        R[a * NB + a] = r;
        for (int b = a + 1; b < nb; ++b) {
          double t = g[a][b];
          for (int c = 0; c < a; ++c) t -= R[c * NB + a] * R[c * NB + b];
          R[a * NB + b] = t / r;
        }
      }
      hs[0] = ok ? 1.0 : 0.0;
      if (NB > 1) hs[1] = minr;
    }
    __syncthreads();
    if (hs[0] == 0.0) return false;
    for (int i = tid; i < n; i += NT) {
      double x[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < nb) {
          double t = Zb[i * NB + c];
          for (int q = 0; q < c; ++q) t -= x[q] * R[q * NB + c];
          x[c] = t / R[c * NB + c];
        } else {
          x[c] = 0.0;
        }
      }
#pragma unroll
      for (int c = 0; c < NB; ++c) Zb[i * NB + c] = x[c];
    }
    __syncthreads();
    return true;
  };

  for (int b0 = c_begin; b0 < c_end; b0 += NB) {
    const int nb = min(NB, c_end - b0);
    for (int e = tid; e < n * NB; e += NT) {
      const int i = e / NB, c = e - i * NB;
      Zb[e] = c < nb ? z[int64_t(i) * k + b0 + c] : 0.0;
    }
    __syncthreads();
    // norms on entry (before any projection) for the dependence test
    if (tid < nb) nrm0[tid] = norms[jb.w_off + b0 + tid];
    __syncthreads();
    if (b0 > c_begin) {
      project(c_begin, b0, 0, nb);
      project(c_begin, b0, 0, nb);
    }
    // fast path: Cholesky QR twice on the sub-block; the column-by-column
    // path below runs only when a column is numerically dependent.  A block
    // that lost most of a column's norm to the earlier blocks had R^-1 scale
    // up what the two projections left of them: project and factor once more.
    bool fast = cholqr(nb, true);
    const bool heavy = fast && NB > 1 && hs[1] < 0.5;
    fast = fast && cholqr(nb, false);
    if (fast && heavy && b0 > 0) {
      project(0, b0, 0, nb);
      fast = cholqr(nb, false);
    }
    if (fast) {
      for (int e = tid; e < n * NB; e += NT) {
        const int i = e / NB, c = e - i * NB;
        if (c < nb) z[int64_t(i) * k + b0 + c] = Zb[e];
      }
      __syncthreads();
      continue;
    }
    // column c of Zb minus its projection on the block's columns q < c (two passes)
    auto inblock = [&](int c) {
      for (int pass = 0; pass < 2 && c > 0; ++pass) {
        double part[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) part[q] = 0;
        for (int i = tid; i < n; i += NT) {
          const double zc = Zb[i * NB + c];
#pragma unroll
          for (int q = 0; q < NB; ++q)
            if (q < c) part[q] += Zb[i * NB + q] * zc;
        }
        block_sums(part, c);
        for (int i = tid; i < n; i += NT) {
          double s = 0;
#pragma unroll
          for (int q = 0; q < NB; ++q)
            if (q < c) s += hs[q] * Zb[i * NB + q];
          Zb[i * NB + c] -= s;
        }
        __syncthreads();
      }
    };
    for (int c = 0; c < nb; ++c) {
      for (int attempt = 0;; ++attempt) {
        inblock(c);
        double part[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) part[q] = 0;
        for (int i = tid; i < n; i += NT) part[0] += Zb[i * NB + c] * Zb[i * NB + c];
        block_sums(part, 1);
        const double nu = sqrt(hs[0]);
        if (nu > 1e-3 * nrm0[c] || attempt == 2) {
          const double f = nu > 0 ? 1.0 / nu : 0.0;
          for (int i = tid; i < n; i += NT) Zb[i * NB + c] *= f;
          __syncthreads();
          if (nu > 0 && nu < 0.5 * nrm0[c]) {
            // most of the column went into the projections, which leave
            // ~eps nrm0 / nu of the earlier columns in it: one more pass
            if (b0 > 0) project(0, b0, c, c + 1);
            inblock(c);
            double p3[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) p3[q] = 0;
            for (int i = tid; i < n; i += NT) p3[0] += Zb[i * NB + c] * Zb[i * NB + c];
            block_sums(p3, 1);
            const double g = hs[0] > 0 ? 1.0 / sqrt(hs[0]) : 0.0;
            for (int i = tid; i < n; i += NT) Zb[i * NB + c] *= g;
            __syncthreads();
          }
          break;
        }
        // numerically dependent: restart from a hashed vector, pulled back into
        // the eigenspace by one solve with this vector's factors when the
        // columns are inverse-iteration vectors (LU != null).  The start is
        // made orthogonal to every earlier column first, so that the solve,
        // which amplifies all eigenvectors of a cluster alike, returns a
        // direction of the cluster outside their span (inverse iteration
        // with reorthogonalisation).
        for (int i = tid; i < n; i += NT)
          Zb[i * NB + c] =
              hash_unit(uint32_t(b0 + c) * 40503u + uint32_t(attempt + 1) * 977u + uint32_t(jb.job), uint32_t(i));
        __syncthreads();
        if (LU != nullptr) {
          if (b0 > 0) project(0, b0, c, c + 1);
          inblock(c);
          if (tid == 0) {
            const double zmax = lu_solve_inplace(LU + jb.lu_off, k, b0 + c, n, Zb + c, NB);
            nrm0[c] = zmax > 0 ? 1.0 / zmax : 1.0;
          }
          __syncthreads();
          const double f = nrm0[c];
          for (int i = tid; i < n; i += NT) Zb[i * NB + c] *= f;
          __syncthreads();
        }
        double p2[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) p2[q] = 0;
        for (int i = tid; i < n; i += NT) p2[0] += Zb[i * NB + c] * Zb[i * NB + c];
        block_sums(p2, 1);
        if (tid == 0) nrm0[c] = sqrt(hs[0]);
        __syncthreads();
        if (b0 > 0) {
          project(0, b0, c, c + 1);
          project(0, b0, c, c + 1);
        }
      }
    }
    for (int e = tid; e < n * NB; e += NT) {
      const int i = e / NB, c = e - i * NB;
      if (c < nb) z[int64_t(i) * k + b0 + c] = Zb[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Lowest eigenpair of a small symmetric matrix, one wavefront.  Lane i owns
// row i.  Tridiagonalisation in LDS, multisection on the Sturm count for the
// smallest eigenvalue (64 points per round), inverse iteration on T by lane 0,
// back-transformation by the stored reflectors.
// Smallest eigenpair of the symmetric tridiagonal (d, e) of order m held in
// LDS, by one wavefront: Gershgorin bracket, multisection on the Sturm count
// (64 points per round), then inverse iteration on T - lambda I (LU with
// partial pivoting, three solves) by lane 0 into z (unit max-norm).  e2 and
// the factor arrays (u0, u1, u2, ml: m each; sw: m) are LDS scratch.
// Returns lambda on every lane.
__device__ double wave_tridiag_lowest(const double* d, const double* e, double* e2, int m, double* z, double* u0,
                                      double* u1, double* u2, double* ml, int* sw) {
  const int l = threadIdx.x & 63;
  double gl = INFINITY, gu = -INFINITY, em = 0;
  for (int i = l; i < m; i += 64) {
    const double el = i > 0 ? fabs(e[i - 1]) : 0.0, er = i + 1 < m ? fabs(e[i]) : 0.0;
    gl = fmin(gl, d[i] - el - er);
    gu = fmax(gu, d[i] + el + er);
    if (i + 1 < m) {
      e2[i] = e[i] * e[i];
      em = fmax(em, e2[i]);
    }
  }
  gl = wave_min(gl);
  gu = wave_max(gu);
  em = wave_max(em);
  __syncthreads();
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = DBL_MIN * fmax(1.0, em);
  double lo = gl - 4.0 * kEps * tnorm - 2.0 * pivmin, hi = gu + 4.0 * kEps * tnorm + 2.0 * pivmin;
  const double atol = 2.0 * kEps * tnorm;
  for (int round = 0; round < 40; ++round) {
    if (hi - lo <= fmax(atol, 2.0 * kEps * fmax(fabs(lo), fabs(hi)))) break;
    const double x = lo + (hi - lo) * double(l + 1) / 65.0;
    const int cnt = (x > lo && x < hi) ? sturm_count(d, e2, m, x, pivmin) : (x >= hi ? 1 : 0);
    const unsigned long long hit = __ballot(cnt >= 1);
    double nlo = lo, nhi = hi;
    if (hit == 0ull) {
      nlo = __shfl(x, 63, 64);
    } else {
      const int f = __ffsll((long long)hit) - 1;
      nhi = __shfl(x, f, 64);
      if (f > 0) nlo = __shfl(x, f - 1, 64);
    }
    if (nlo == lo && nhi == hi) break;
    lo = nlo;
    hi = nhi;
  }
  const double lam = 0.5 * (lo + hi);
  if (l == 0) {
    if (m == 1) {
      z[0] = 1.0;
    } else {
      const double piv = kEps * fmax(tnorm, DBL_MIN);
      auto guard = [&](double x) { return fabs(x) < piv ? (x < 0 ? -piv : piv) : x; };
      double c0 = d[0] - lam, c1 = e[0];
      for (int i = 0; i + 1 < m; ++i) {
        const double ei = e[i], dn = d[i + 1] - lam, en = i + 2 < m ? e[i + 1] : 0.0;
        if (fabs(c0) >= fabs(ei)) {
          c0 = guard(c0);
          const double mu = ei / c0;
          u0[i] = c0;
          u1[i] = c1;
          u2[i] = 0.0;
          ml[i] = mu;
          sw[i] = 0;
          c0 = dn - mu * c1;
          c1 = en;
        } else {
          const double mu = c0 / ei;
          u0[i] = ei;
          u1[i] = dn;
          u2[i] = en;
          ml[i] = mu;
          sw[i] = 1;
          c0 = c1 - mu * dn;
          c1 = -mu * en;
        }
      }
      u0[m - 1] = guard(c0);
      u1[m - 1] = 0.0;
      u2[m - 1] = 0.0;
      for (int i = 0; i < m; ++i) z[i] = 1.0 + 0.01 * double((i * 37) % 11);
      for (int itn = 0; itn < 3; ++itn) {
        double cur = z[0];
        for (int i = 0; i + 1 < m; ++i) {
          double nxt = z[i + 1];
          if (sw[i]) {
            const double t = cur;
            cur = nxt;
            nxt = t;
          }
          nxt -= ml[i] * cur;
          z[i] = cur;
          cur = nxt;
        }
        z[m - 1] = cur;
        double z1 = 0, z2 = 0, zmax = 0;
        for (int i = m - 1; i >= 0; --i) {
          const double zi = (z[i] - u1[i] * z1 - u2[i] * z2) / u0[i];
          z[i] = zi;
          zmax = fmax(zmax, fabs(zi));
          z2 = z1;
          z1 = zi;
        }
        const double s = zmax > 0 ? 1.0 / zmax : 1.0;
        for (int i = 0; i < m; ++i) z[i] *= s;
      }
    }
  }
  __syncthreads();
  return lam;
}

__global__ __launch_bounds__(64) void lowest_eig_kernel(const double* __restrict__ H, int ld, int m,
                                                        double* __restrict__ c_out, double* __restrict__ theta) {
  __shared__ double a[64][65];  // full symmetric matrix; column k below the diagonal then holds reflector k
  __shared__ double tau[64], d[64], e[64], e2[64], vv[64], ww[64], z[64];
  __shared__ double u0[64], u1[64], u2[64], ml[64];
  __shared__ int sw[64];
  const int l = threadIdx.x;
  for (int i = 0; i < m; ++i)
    if (l <= i) {
      const double x = H[int64_t(i) * ld + l];
      a[i][l] = x;
      a[l][i] = x;
    }
  __syncthreads();
  for (int k = 0; k + 2 < m; ++k) {
    const double x = (l > k && l < m) ? a[l][k] : 0.0;
    const double alpha = a[k + 1][k];
    const double dk = a[k][k];
    const double sigma = wave_sum(l >= k + 2 && l < m ? x * x : 0.0);
    double tk = 0, beta = alpha, scale = 0;
    if (sigma > 0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0 ? -nrm : nrm;
      tk = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    const double vi = l == k + 1 ? 1.0 : (l >= k + 2 && l < m ? x * scale : 0.0);
    __syncthreads();
    if (l > k && l < m) a[l][k] = vi;
    vv[l] = vi;
    if (l == 0) {
      d[k] = dk;
      e[k] = beta;
      tau[k] = tk;
    }
    __syncthreads();
    double p = 0;
    if (l > k && l < m)
      for (int j = k + 1; j < m; ++j) p += a[l][j] * vv[j];
    const double pv = wave_sum(p * vi);
    const double wi = tk * p - 0.5 * tk * tk * pv * vi;
    ww[l] = wi;
    __syncthreads();
    if (l > k && l < m)
      for (int j = k + 1; j < m; ++j) a[l][j] -= vi * ww[j] + wi * vv[j];
    __syncthreads();
  }
  if (l == 0) {
    if (m >= 2) {
      d[m - 2] = a[m - 2][m - 2];
      e[m - 2] = a[m - 1][m - 2];
      tau[m - 2] = 0;
    }
    d[m - 1] = a[m - 1][m - 1];
    e[m - 1] = 0;
    tau[m - 1] = 0;
  }
  __syncthreads();
  const double lam = wave_tridiag_lowest(d, e, e2, m, z, u0, u1, u2, ml, sw);
  // back-transformation y = H_0 ... H_{m-3} z
  double y = l < m ? z[l] : 0.0;
  for (int k = m - 3; k >= 0; --k) {
    const double vi = (l > k && l < m) ? a[l][k] : 0.0;
    const double s = wave_sum(vi * y);
    y -= tau[k] * s * vi;
  }
  const double nrm = sqrt(wave_sum(y * y));
  if (l < m) c_out[l] = y / nrm;
  if (l == 0) *theta = lam;
}

// Lowest eigenpair of the Lanczos tridiagonal T_m (alpha, beta), m <= kTriMax:
// z (m) normalised, *theta, and the residual estimate |beta_next z[m-1]| of
// the Ritz pair (beta_next: the norm of the next Lanczos vector).
constexpr int kTriMax = 512;
__global__ __launch_bounds__(64) void tridiag_lowest_kernel(const double* __restrict__ alpha,
                                                            const double* __restrict__ beta, int m,
                                                            const double* __restrict__ beta_next,
                                                            double* __restrict__ z_out, double* __restrict__ theta,
                                                            double* __restrict__ est) {
  __shared__ double d[kTriMax], e[kTriMax], e2[kTriMax], z[kTriMax];
  __shared__ double u0[kTriMax], u1[kTriMax], u2[kTriMax], ml[kTriMax];
  __shared__ int sw[kTriMax];
  const int l = threadIdx.x;
  for (int i = l; i < m; i += 64) {
    d[i] = alpha[i];
    e[i] = i + 1 < m ? beta[i] : 0.0;
  }
  __syncthreads();
  const double lam = wave_tridiag_lowest(d, e, e2, m, z, u0, u1, u2, ml, sw);
  double s = 0;
  for (int i = l; i < m; i += 64) s += z[i] * z[i];
  const double nrm = sqrt(wave_sum(s));
  for (int i = l; i < m; i += 64) z_out[i] = z[i] / nrm;
  if (l == 0) {
    *theta = lam;
    if (est) *est = fabs(*beta_next * z[m - 1] / nrm);
  }
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    TK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(bytes)));
}

template <int NT, int T, bool kLds>
void launch_tridiag_t(Context* ctx, const TailJob* jobs, int n_jobs, int mmax, double* A, double* D, double* E,
                      double* TAU, const double* VP, int64_t w_n) {
  constexpr int NW = NT / 64;
  const size_t lds =
      ((kLds ? size_t(mmax) * mmax : 0) + (4 + NW) * size_t(mmax) + 2 * NW + 8) * sizeof(double);
  allow_lds(tridiag_kernel<NT, T, kLds>, lds);
  hipLaunchKernelGGL((tridiag_kernel<NT, T, kLds>), dim3(n_jobs), dim3(NT), lds, ctx->stream, jobs, A, D, E, TAU,
                     VP, w_n);
}

constexpr int kRegTailNT = 512, kRegTail = 192;

void launch_tridiag_reg(Context* ctx, const TailJob* jobs, int n_jobs, double* A, double* D, double* E, double* TAU,
                        const double* VP, int64_t w_n) {
  constexpr int NW = kRegTailNT / 64, RPW = kRegTail / NW;
  const size_t lds = (4 * size_t(kRegTail) + 2 * NW + 8 + size_t(NW) * 64 * (RPW + 1)) * sizeof(double);
  allow_lds(tridiag_reg_kernel<kRegTailNT, kRegTail>, lds);
  hipLaunchKernelGGL((tridiag_reg_kernel<kRegTailNT, kRegTail>), dim3(n_jobs), dim3(kRegTailNT), lds, ctx->stream,
                     jobs, A, D, E, TAU, VP, w_n);
}

// Trailing size below which a job's tridiagonalisation finishes in one
// workgroup: 192 with the block in registers (TK_TRIDIAG_TAIL = 64, 96 or 128
// selects the LDS-resident kernel at that size, for comparisons).
int tridiag_tail() {
  static const int t = [] {
    const char* e = std::getenv("TK_TRIDIAG_TAIL");
    const int v = e ? std::atoi(e) : 0;
    return (v == 64 || v == 96 || v == 128) ? v : kRegTail;
  }();
  return t;
}

// Algorithmic work of the vector phase: CGS2 reads the earlier columns twice
// per block of NB columns (n k^2 / NB doubles) and does 4 n k^2 flops; the
// back-transformation reads each reflector once per column block and
// updates n k values per reflector (2 n^2 k flops).
// ---------------------------------------------------------------------------
// 3d. Compact-WY back-transformation.  Reflectors kk = j0 .. j0+nb-1 form
// H_j0 ... H_(j0+nb-1) = I - V T V^T (V: n x nb, v_kk[kk+1] = 1, zero above;
// T upper triangular, T[l][l] = tau_l, T[l][i] = -tau_i sum_m T[l][m] S[m][i]
// with S = V^T V).  Rows of A become the rows of V^T in place.
constexpr int kWyNb = 128;   // reflectors per block
constexpr int kWySeg = 256;  // rows per split-K segment of V^T Y

// Row kk of every job: zero columns <= kk, 1 at kk+1 (rows n-2, n-1: zero).
__global__ __launch_bounds__(256) void wy_rows_kernel(const DevJob* __restrict__ jobs, double* __restrict__ A) {
  const DevJob jb = jobs[blockIdx.y];
  const int n = jb.n;
  double* a = A + jb.a_off;
  for (int kk = blockIdx.x; kk < n; kk += gridDim.x) {
    double* row = a + int64_t(kk) * n;
    const int lim = kk + 2 < n ? kk + 1 : n;  // columns [0, lim) cleared
    for (int j = threadIdx.x; j < lim; j += blockDim.x) row[j] = 0.0;
    if (threadIdx.x == 0 && kk + 2 < n) row[kk + 1] = 1.0;
    if (kk + 2 >= n)
      for (int j = lim + threadIdx.x; j < n; j += blockDim.x) row[j] = 0.0;
  }
}

struct WyBlock {
  int64_t t_off;    // S and T (kWyNb x kWyNb, ld kWyNb) of this block
  int64_t tau_off;  // tau of reflector j0
  int32_t nb, pad;
};

// T column by column in LDS: T[0:i, i] = -tau_i T[0:i, 0:i] S[0:i, i].  Four
// lanes per row l (lane q sums the terms m = l + q, l + q + 4, ...; the quad
// is closed by two butterfly steps in a fixed order), the column S[:, i]
// staged in LDS; the next column's global loads are issued before this
// column's sums.
constexpr int kWyTLanes = 4;
constexpr size_t kWyTLds = (size_t(kWyNb) * (kWyNb + 1) + 2 * kWyNb) * sizeof(double);

__global__ __launch_bounds__(kWyNb * kWyTLanes) void wy_t_kernel(const WyBlock* __restrict__ blocks,
                                                                 const double* __restrict__ S,
                                                                 const double* __restrict__ TAUg,
                                                                 double* __restrict__ T) {
  extern __shared__ double sm[];
  double* tl = sm;                        // kWyNb x (kWyNb + 1)
  double* scb = sm + kWyNb * (kWyNb + 1);  // 2 x kWyNb: column i of S, double-buffered
  const WyBlock b = blocks[blockIdx.x];
  const double* s = S + b.t_off;
  const double* tau = TAUg + b.tau_off;
  const int l = threadIdx.x / kWyTLanes, q = threadIdx.x % kWyTLanes, nb = b.nb;
  for (int m = q; m < kWyNb; m += kWyTLanes) tl[l * (kWyNb + 1) + m] = 0.0;
  double snext = 0.0;
  // Row l of T is written and read only by its own kWyTLanes lanes (one
  // wavefront), so the one barrier per step publishes the shared column of
  // S; the column alternates between two buffers, so that the next step's
  // store cannot meet a read of this one.
  for (int i = 0; i < nb; ++i) {
    double* sc = scb + (i & 1) * kWyNb;
    if (q == 0) sc[l] = snext;
    snext = (q == 0 && l < i + 1 && i + 1 < nb) ? s[int64_t(l) * kWyNb + i + 1] : 0.0;
    __syncthreads();
    // every lane takes part in the butterfly (uniform control flow); four
    // partial sums keep a lane's loads and products independent
    double a = 0;
    if (l < i) {
      const double* tr = tl + l * (kWyNb + 1);
      double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
      int m = l + q;
      for (; m + 3 * kWyTLanes < i; m += 4 * kWyTLanes) {
        a0 += tr[m] * sc[m];
        a1 += tr[m + kWyTLanes] * sc[m + kWyTLanes];
        a2 += tr[m + 2 * kWyTLanes] * sc[m + 2 * kWyTLanes];
        a3 += tr[m + 3 * kWyTLanes] * sc[m + 3 * kWyTLanes];
      }
      for (; m < i; m += kWyTLanes) a0 += tr[m] * sc[m];
      a = (a0 + a1) + (a2 + a3);
    }
    a += __shfl_xor(a, 1, 64);
    a += __shfl_xor(a, 2, 64);
    if (q == 0) {
      if (l < i) tl[l * (kWyNb + 1) + i] = -tau[i] * a;
      else if (l == i) tl[l * (kWyNb + 1) + i] = tau[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  double* t = T + b.t_off;
  for (int m = q; m < kWyNb; m += kWyTLanes) t[int64_t(l) * kWyNb + m] = tl[l * (kWyNb + 1) + m];
}

struct VecWork {
  double orth_bytes = 0, orth_flops = 0, bt_bytes = 0, bt_flops = 0;
};

// Z (output, ld n, the first k columns at a_off) = Q * zi, with Q the product
// of the reflectors stored in the rows of A; blocks of kWyNb reflectors are
// applied last block first, each as W = V^T Y (split over row segments) and
// Y -= V T W (one contraction summing the segments).
void wy_backtransform(Context* ctx, const std::vector<DevJob>& dj, const DevJob* djd, double* A, int64_t a_n,
                      const double* TAU, const double* zi, int64_t z_tot, double* Z, int64_t z_n) {
  const int nj = int(dj.size());
  int nmax = 0;
  for (const DevJob& j : dj) nmax = std::max(nmax, j.n);
  // Y = zi (ld k) into Z (ld n)
  std::vector<CopyJob> cp;
  for (const DevJob& j : dj) cp.push_back({j.z_off, j.a_off, j.n, j.k, j.k, j.n, 0, 0});
  launch_copies(ctx, cp, zi, z_tot, Z, z_n);
  if (nmax < 3) return;
  hipLaunchKernelGGL(wy_rows_kernel, dim3(unsigned(std::min(nmax, 1024)), unsigned(nj)), dim3(256), 0, ctx->stream,
                     djd, A);
  TK_HIP(hipGetLastError());
  // blocks of every job
  struct HB {
    int job, j0, nb;
  };
  std::vector<HB> hb;
  std::vector<WyBlock> wb;
  std::vector<std::vector<int>> blocks_of(nj);
  for (int q = 0; q < nj; ++q) {
    const int n = dj[q].n;
    for (int j0 = 0; j0 + 2 < n; j0 += kWyNb) {
      const int nb = std::min(kWyNb, n - 2 - j0);
      blocks_of[q].push_back(int(hb.size()));
      wb.push_back({int64_t(hb.size()) * kWyNb * kWyNb, dj[q].w_off + j0, nb, 0});
      hb.push_back({q, j0, nb});
    }
  }
  if (hb.empty()) return;
  const int64_t st_n = int64_t(hb.size()) * kWyNb * kWyNb;
  DBuf<double> S(ctx, size_t(st_n)), T(ctx, size_t(st_n)), Ut(ctx, size_t(a_n));
  DBuf<WyBlock> wbd(ctx, wb.size());
  wbd.upload(wb.data(), wb.size());
  // every list of the back-transformation, uploaded with one copy
  ListPack pack;
  int i_s, i_ut;
  {  // S = V^T V per block
    ContractList cl;
    cl.flags = kTransB;
    for (size_t b = 0; b < hb.size(); ++b) {
      const DevJob& j = dj[hb[b].job];
      const int j0 = hb[b].j0, nb = hb[b].nb;
      const int64_t v = j.a_off + int64_t(j0) * j.n + (j0 + 1);
      cl.begin_out(wb[b].t_off, nb, nb, kWyNb);
      cl.begin_term(v, j.n, j.n - j0 - 1);
      cl.add_src(v, j.n, 1.0);
    }
    i_s = pack.add(std::move(cl), a_n, a_n, st_n);
  }
  {  // U^T = T^T V^T per block, stored in the rows of Ut like V^T in A
    ContractList cl;
    cl.flags = kTransA;
    for (size_t b = 0; b < hb.size(); ++b) {
      const DevJob& j = dj[hb[b].job];
      const int j0 = hb[b].j0, nb = hb[b].nb;
      const int64_t v = j.a_off + int64_t(j0) * j.n + (j0 + 1);
      cl.begin_out(v, nb, j.n - j0 - 1, j.n);
      cl.begin_term(v, j.n, nb);
      cl.add_src(wb[b].t_off, kWyNb, 1.0);
    }
    i_ut = pack.add(std::move(cl), st_n, a_n, a_n);
  }
  // application, last block of every job first
  size_t steps = 0;
  int64_t w_max = 0;
  for (int q = 0; q < nj; ++q) steps = std::max(steps, blocks_of[q].size());
  for (size_t st = 0; st < steps; ++st) {
    int64_t w = 0;
    for (int q = 0; q < nj; ++q) {
      const size_t nbq = blocks_of[q].size();
      if (st >= nbq) continue;
      const HB& h = hb[blocks_of[q][nbq - 1 - st]];
      const int rows = dj[q].n - h.j0 - 1;
      w += int64_t((rows + kWySeg - 1) / kWySeg) * h.nb * dj[q].k;
    }
    w_max = std::max(w_max, w);
  }
  DBuf<double> Wb(ctx, size_t(std::max<int64_t>(w_max, 1)));
  std::vector<std::pair<int, int>> step_lists(steps, {-1, -1});
  for (size_t st = 0; st < steps; ++st) {
    ContractList lw, lu;
    lw.flags = 0;
    lu.flags = kTransA | kTransC;
    lu.beta = 1.0;
    int64_t w = 0;
    for (int q = 0; q < nj; ++q) {
      const size_t nbq = blocks_of[q].size();
      if (st >= nbq) continue;
      const DevJob& j = dj[q];
      const HB& h = hb[blocks_of[q][nbq - 1 - st]];
      const int n = j.n, k = j.k, j0 = h.j0, nb = h.nb;
      const int64_t vrow = j.a_off + int64_t(j0) * n;
      // Y^T (rows j0+1..n-1 of Y) -= (sum_seg W_seg)^T U^T
      lu.begin_out(j.a_off + int64_t(j0 + 1) * n, k, n - j0 - 1, n);
      lu.begin_term(vrow + j0 + 1, n, nb);
      for (int r0 = j0 + 1; r0 < n; r0 += kWySeg) {
        const int r1 = std::min(n, r0 + kWySeg);
        lw.begin_out(w, nb, k, k);
        lw.begin_term(j.a_off + int64_t(r0) * n, n, r1 - r0);
        lw.add_src(vrow + r0, n, 1.0);
        lu.add_src(w, k, -1.0);
        w += int64_t(nb) * k;
      }
    }
    step_lists[st].first = pack.add(std::move(lw), a_n, z_n, w_max);
    step_lists[st].second = pack.add(std::move(lu), w_max, a_n, z_n);
  }
  pack.upload(ctx);
  auto run = [&](int i, const double* A_, const double* B_, double* C_) {
    if (i >= 0) launch_contract(ctx, pack[i], A_, B_, C_);
  };
  run(i_s, A, A, S.get());
  allow_lds(wy_t_kernel, kWyTLds);
  hipLaunchKernelGGL(wy_t_kernel, dim3(unsigned(hb.size())), dim3(kWyNb * kWyTLanes), kWyTLds, ctx->stream, wbd.get(),
                     S.get(),
                     TAU, T.get());
  TK_HIP(hipGetLastError());
  run(i_ut, T.get(), A, Ut.get());
  for (size_t st = 0; st < steps; ++st) {
    run(step_lists[st].first, A, Z, Wb.get());
    run(step_lists[st].second, Wb.get(), Ut.get(), Z);
  }
}

// Rows per partial Gram block of the block orthonormalisation: G = Z_b^T Z_b
// over n rows as ceil(n / kGramSeg) independent contractions of kGramSeg
// rows (short inner chains), summed by cholinv_kernel when it loads G.
constexpr int kGramSeg = 256;

// Cholesky factor and inverse of the nb x nb Gram block G = Z_b^T Z_b of each
// job (one workgroup per job, goff[job] < 0: no block): G = R^T R and
// Rinv = R^-1 (upper triangular, nb x nb at goff).  With `test`, the block is
// left to the column-by-column path when some R_cc <= 0.1 of its column's
// norm on entry (norms at w_off + c0 + c): R^-1 would scale what the two
// projections left of the earlier columns up past ~10 eps.  A job with such
// a column or a non-positive pivot, or
// (without `test`) one whose ok flag is already 0, gets Rinv = I and ok = 0.
template <int NB>
__global__ __launch_bounds__(256) void cholinv_kernel(const DevJob* __restrict__ jobs, const int64_t* __restrict__ goff,
                                                      const double* __restrict__ G, double* __restrict__ Rinv,
                                                      const double* __restrict__ norms, int step, int c_count,
                                                      int test, int* __restrict__ ok) {
  extern __shared__ double sm[];
  double* L = sm;                 // NB x (NB + 1): G, then the rows of R unscaled
  double* W = sm + NB * (NB + 1);  // NB x (NB + 1): R^-T with its rows unscaled
  __shared__ int good;
  __shared__ double nrm[NB];
  const int64_t off = goff[blockIdx.x];
  if (off < 0) return;
  const DevJob jb = jobs[blockIdx.x];
  const int c0 = jb.pad + step * c_count;
  const int nb = min(c_count, jb.k - c0);
  const int tid = threadIdx.x;
  constexpr int LD = NB + 1;
  // 16 x 16 thread grid over the block (no index division in the loops); G
  // arrives as one partial Gram block per segment of kGramSeg rows, summed
  // here in segment order
  // (the 16 elements of a thread are loaded together per segment: a
  // per-element loop over the segments would wait on every load in turn)
  const int tx = tid & 15, ty = tid >> 4;
  const int nseg = (jb.n + kGramSeg - 1) / kGramSeg;
  static_assert(NB <= 64, "four rows and columns per thread");
  double gs[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) gs[a][b] = 0.0;
  for (int sg = 0; sg < nseg; ++sg) {
    const double* Gs = G + off + int64_t(sg) * nb * nb;
    double v[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int i = ty + 16 * a, j = tx + 16 * b;
        v[a][b] = (i < nb && j < nb) ? Gs[int64_t(i) * nb + j] : 0.0;
      }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) gs[a][b] += v[a][b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = ty + 16 * a, j = tx + 16 * b;
      if (i < nb && j < nb) L[i * LD + j] = gs[a][b];
    }
  if (test && tid < nb) nrm[tid] = norms[jb.w_off + c0 + tid];
  if (tid == 0) good = test ? 1 : ok[blockIdx.x];
  __syncthreads();
  // Right-looking Cholesky G = R^T R with R^-1 formed in the same steps, one
  // barrier per column.  Y = R^-T (lower triangular) solves R^T Y = I by
  // rows: Y_a = (e_a - sum_{l<a} R_la Y_l) / R_aa.  W starts as I; step a
  // reads the pivot d = R_aa^2 (every thread the same value, so the same
  // decision) and, with u_b = L_ab / d for the rows b > a below it, updates
  // the trailing block L_bc -= u_b L_ac (c >= b) and W_bc -= u_b W_ac (c <= a)
  // with row a unscaled: R_ab Y_a = u_b W_a.  Row a of W is final then;
  // Y_a = W_a / R_aa and R^-1 = Y^T, scaled by the stored pivots on the way
  // out.  Row a is read and only rows below it written within a step.
  __shared__ double piv[NB];
  for (int i = ty; i < nb; i += 16)
    for (int j = tx; j < nb; j += 16) W[i * LD + j] = i == j ? 1.0 : 0.0;
  __syncthreads();
  bool gl = good != 0;
  for (int a = 0; a < nb && gl; ++a) {
    const double dd = L[a * LD + a];
    if (!(dd > 0) || (test && sqrt(dd) <= 0.1 * nrm[a])) {
      gl = false;
      break;
    }
    if (tid == 0) piv[a] = sqrt(dd);
    const double idd = 1.0 / dd;
    // rows b = a + 1 + ty + 16 r, columns c = tx + 16 s: entry (b, c) is the
    // trailing block's when c >= b and Y's otherwise (W_a is zero beyond
    // column a, so Y's entries a < c < b stay as they are); chosen by address,
    // so that a thread's loads are issued together rather than one per branch
    double ub[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) ub[r] = L[a * LD + min(a + 1 + ty + 16 * r, NB - 1)] * idd;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int b = a + 1 + ty + 16 * r;
      const int bl = min(b, NB - 1);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = tx + 16 * q;
        double* M = c >= b ? L : W;
        const double v = M[bl * LD + c] - ub[r] * M[a * LD + c];
        if (b < nb && c < nb) M[bl * LD + c] = v;
      }
    }
    __syncthreads();
  }
  const bool g = gl;
  __syncthreads();
  for (int i = ty; i < nb; i += 16)
    for (int j = tx; j < nb; j += 16)
      Rinv[off + int64_t(i) * nb + j] = g ? (i <= j ? W[j * LD + i] / piv[j] : 0.0) : (i == j ? 1.0 : 0.0);
  if (tid == 0) ok[blockIdx.x] = g ? 1 : 0;
}

constexpr int kOrthSeg = 256;  // rows per split-K segment of Z^T Z_block
constexpr int kCholNB = 64;    // largest block finished by Cholesky QR

// Columns per block of the projection contractions (TK_ORTH_BLOCK overrides, for tuning).
int orth_block() {
  static const int b = [] {
    const char* e = std::getenv("TK_ORTH_BLOCK");
    const int v = e ? std::atoi(e) : 0;
    return (v >= 8 && v <= 256 && v % 8 == 0) ? v : 64;
  }();
  return b;
}

template <int NB>
void launch_orth_range(Context* ctx, int nmax, int nj, const DevJob* djd, double* Z, const double* lu, int step,
                       int c_count, const double* norms, const int* done) {
  constexpr int kNT = 512, NW = kNT / 64;
  const size_t lds = (size_t(nmax) * NB + NW * 64 * NB + 64 * NB + NW * NB + 2 * NB) * sizeof(double);
  allow_lds(orth_kernel<kNT, NB>, lds);
  hipLaunchKernelGGL((orth_kernel<kNT, NB>), dim3(nj), dim3(kNT), lds, ctx->stream, djd, Z, lu, step, c_count,
                     norms, done);
  TK_HIP(hipGetLastError());
}

// Blocked classical Gram-Schmidt applied twice: block b of kOrthBlock columns
// is projected against all earlier columns by two grouped contractions per
// pass (H = Z^T Z_b split over row segments, Z_b -= Z H summing the
// segments), then orthonormalised within itself by orth_kernel.  Norms on
// entry come from one pass over Z; w_off of each job indexes its norms.
void orth_blocked(Context* ctx, const std::vector<DevJob>& dj, const DevJob* djd, double* Z, int64_t z_n,
                  const double* lu, int64_t w_n) {
  const int nj = int(dj.size());
  int kmax = 0, nmax = 0;
  for (const DevJob& j : dj) {
    kmax = std::max(kmax, j.k);
    nmax = std::max(nmax, j.n);
  }
  if (kmax == 0) return;
  const int kOrthBlock = orth_block();
  DBuf<double> norms(ctx, size_t(std::max<int64_t>(w_n, 1)));
  hipLaunchKernelGGL(colnorm_kernel, dim3(unsigned((kmax + 15) / 16), unsigned(nj)), dim3(256), 0, ctx->stream,
                     djd, Z, norms.get());
  TK_HIP(hipGetLastError());
  // block `step` of job j covers columns [first_j + step * kOrthBlock, ...)
  int steps = 0;
  for (const DevJob& j : dj) steps = std::max(steps, (j.k - j.pad + kOrthBlock - 1) / kOrthBlock);
  int64_t h_max = 1;
  for (int st = 0; st < steps; ++st) {
    int64_t h = 0;
    for (const DevJob& j : dj) {
      const int c0 = j.pad + st * kOrthBlock;
      if (c0 > 0 && c0 < j.k) h += int64_t((j.n + kOrthSeg - 1) / kOrthSeg) * c0 * std::min(kOrthBlock, j.k - c0);
    }
    h_max = std::max(h_max, h);
  }
  DBuf<double> H(ctx, size_t(h_max));
  // Cholesky QR scratch sized for the largest step: Gram blocks and R^-1
  // (nb x nb per job), T (n x nb per job); per step and job the block offset
  // (-1: none) and the success flag
  int64_t cq_gt = 1, cq_tt = 1;
  std::vector<int64_t> cq_offs(size_t(steps) * nj, -1);
  for (int st = 0; st < steps; ++st) {
    int64_t gt = 0, tt = 0;
    for (int o = 0; o < nj; ++o) {
      const DevJob& j = dj[size_t(o)];
      const int c0 = j.pad + st * kOrthBlock;
      if (c0 >= j.k) continue;
      const int nb = std::min(kOrthBlock, j.k - c0);
      cq_offs[size_t(st) * nj + o] = gt;
      gt += int64_t((j.n + kGramSeg - 1) / kGramSeg) * nb * nb;
      tt += int64_t(j.n) * nb;
    }
    cq_gt = std::max(cq_gt, gt);
    cq_tt = std::max(cq_tt, tt);
  }
  DBuf<double> cq_g(ctx, size_t(cq_gt)), cq_r(ctx, size_t(cq_gt)), cq_t(ctx, size_t(cq_tt));
  DBuf<int64_t> cq_off(ctx, std::max<size_t>(cq_offs.size(), 1));
  if (!cq_offs.empty()) cq_off.upload(cq_offs.data(), cq_offs.size());
  DBuf<int> cq_ok(ctx, std::max<size_t>(size_t(steps) * nj, 1));
  // every step's lists first, uploaded with one copy; then the launches
  struct StepLists {
    int h[2] = {-1, -1}, u2[2] = {-1, -1};
    int g = -1, t = -1, g2 = -1, z = -1;
    bool chol = false;
  };
  std::vector<StepLists> sl(static_cast<size_t>(steps));
  ListPack pack;
  for (int st = 0; st < steps; ++st) {
    for (int pass = 0; pass < 2; ++pass) {
      ContractList lh, lu2;
      lh.flags = kTransA;
      lu2.flags = kTransA | kTransB | kTransC;
      lu2.beta = 1.0;
      int64_t h = 0;
      for (const DevJob& j : dj) {
        const int c0 = j.pad + st * kOrthBlock;
        if (c0 <= 0 || c0 >= j.k) continue;
        const int nb = std::min(kOrthBlock, j.k - c0);
        // Z_b^T (nb x n) -= (sum_seg H_seg)^T Z[:, 0:c0]^T
        lu2.begin_out(j.z_off + c0, nb, j.n, j.k);
        lu2.begin_term(j.z_off, j.k, c0);
        for (int r0 = 0; r0 < j.n; r0 += kOrthSeg) {
          const int r1 = std::min(j.n, r0 + kOrthSeg);
          // H_seg (c0 x nb) = Z[r0:r1, 0:c0]^T Z[r0:r1, c0:c0+nb]
          lh.begin_out(h, c0, nb, nb);
          lh.begin_term(j.z_off + int64_t(r0) * j.k + c0, j.k, r1 - r0);
          lh.add_src(j.z_off + int64_t(r0) * j.k, j.k, 1.0);
          lu2.add_src(h, nb, -1.0);
          h += int64_t(c0) * nb;
        }
      }
      sl[size_t(st)].h[pass] = pack.add(std::move(lh), z_n, z_n, h_max);
      sl[size_t(st)].u2[pass] = pack.add(std::move(lu2), h_max, z_n, z_n);
    }
    // within the block: Cholesky QR twice through grouped contractions,
    // G = Z_b^T Z_b = R^T R, T = Z_b R^-1, then the same from T back into
    // Z_b; a block with a numerically dependent column keeps Z_b (R^-1 = I)
    // and is finished column by column by orth_kernel
    if (kOrthBlock <= kCholNB) {
      int64_t gt = 0, tt = 0;
      ContractList lg, lt, lg2, lz;
      lg.flags = kTransA;
      lg2.flags = kTransA;
      for (int o = 0; o < nj; ++o) {
        const DevJob& j = dj[size_t(o)];
        const int c0 = j.pad + st * kOrthBlock;
        if (c0 >= j.k) continue;
        const int nb = std::min(kOrthBlock, j.k - c0);
        // partial Gram blocks per kGramSeg rows (summed by cholinv_kernel)
        for (int r0 = 0, sg = 0; r0 < j.n; r0 += kGramSeg, ++sg) {
          const int rows = std::min(kGramSeg, j.n - r0);
          const int64_t gs = gt + int64_t(sg) * nb * nb;
          lg.begin_out(gs, nb, nb, nb);
          lg.begin_term(j.z_off + int64_t(r0) * j.k + c0, j.k, rows);
          lg.add_src(j.z_off + int64_t(r0) * j.k + c0, j.k, 1.0);
          lg2.begin_out(gs, nb, nb, nb);
          lg2.begin_term(tt + int64_t(r0) * nb, nb, rows);
          lg2.add_src(tt + int64_t(r0) * nb, nb, 1.0);
        }
        lt.begin_out(tt, j.n, nb, nb);
        lt.begin_term(gt, nb, nb);
        lt.add_src(j.z_off + c0, j.k, 1.0);
        lz.begin_out(j.z_off + c0, j.n, nb, j.k);
        lz.begin_term(gt, nb, nb);
        lz.add_src(tt, nb, 1.0);
        gt += int64_t((j.n + kGramSeg - 1) / kGramSeg) * nb * nb;
        tt += int64_t(j.n) * nb;
      }
      if (gt > 0) {
        StepLists& q = sl[size_t(st)];
        q.chol = true;
        q.g = pack.add(std::move(lg), z_n, z_n, cq_gt);
        q.t = pack.add(std::move(lt), z_n, cq_gt, cq_tt);
        q.g2 = pack.add(std::move(lg2), cq_tt, cq_tt, cq_gt);
        q.z = pack.add(std::move(lz), cq_tt, cq_gt, z_n);
      }
    }
  }
  pack.upload(ctx);
  auto run = [&](int i, const double* A, const double* B, double* C) {
    if (i >= 0) launch_contract(ctx, pack[i], A, B, C);
  };
  const size_t chol_lds = 2 * size_t(kCholNB) * (kCholNB + 1) * sizeof(double);
  if (kOrthBlock <= kCholNB) allow_lds(cholinv_kernel<kCholNB>, chol_lds);
  for (int st = 0; st < steps; ++st) {
    const StepLists& q = sl[size_t(st)];
    for (int pass = 0; pass < 2; ++pass) {
      run(q.h[pass], Z, Z, H.get());
      run(q.u2[pass], H.get(), Z, Z);
    }
    const int* done = nullptr;
    if (q.chol) {
      const int64_t* offs = cq_off.get() + int64_t(st) * nj;
      int* okf = cq_ok.get() + int64_t(st) * nj;
      run(q.g, Z, Z, cq_g.get());
      hipLaunchKernelGGL(cholinv_kernel<kCholNB>, dim3(unsigned(nj)), dim3(256), chol_lds, ctx->stream, djd, offs,
                         cq_g.get(), cq_r.get(), norms.get(), st, kOrthBlock, 1, okf);
      run(q.t, Z, cq_r.get(), cq_t.get());
      run(q.g2, cq_t.get(), cq_t.get(), cq_g.get());
      hipLaunchKernelGGL(cholinv_kernel<kCholNB>, dim3(unsigned(nj)), dim3(256), chol_lds, ctx->stream, djd, offs,
                         cq_g.get(), cq_r.get(), norms.get(), st, kOrthBlock, 0, okf);
      run(q.z, cq_t.get(), cq_r.get(), Z);
      TK_HIP(hipGetLastError());
      done = okf;
    }
    if (nmax <= 1024)
      launch_orth_range<8>(ctx, nmax, nj, djd, Z, lu, st, kOrthBlock, norms.get(), done);
    else
      launch_orth_range<4>(ctx, nmax, nj, djd, Z, lu, st, kOrthBlock, norms.get(), done);
  }
}

}  // namespace

void SymEigBatch::values(const std::vector<SymJob>& jobs, double* A, int64_t a_n, double* W, int64_t w_n,
                         bool descending) {
  jobs_ = jobs;
  A_ = A;
  W_ = W;
  a_n_ = a_n;
  nmax_ = 0;
  for (const SymJob& j : jobs) {
    if (j.n < 1 || j.n > kSymEigMaxN) throw ArgError("eigensolver: matrix dimension outside 1..2048");
    if (j.a_off < 0 || j.a_off + int64_t(j.n) * j.n > a_n || j.w_off < 0 || j.w_off + j.n > w_n)
      throw std::out_of_range("eigensolver: job outside its buffers");
    nmax_ = std::max(nmax_, j.n);
  }
  if (jobs.empty()) return;
  // d, e, tau share the eigenvalue offsets
  de_.reset(ctx_, size_t(3 * std::max<int64_t>(w_n, 1)));
  tn_.reset(ctx_, jobs.size());
  double* D = de_.get();
  double* E = D + w_n;
  double* TAU = E + w_n;
  // tridiagonalisation, largest matrices first
  std::vector<int> order(jobs.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return jobs[x].n > jobs[y].n; });
  // column steps of job q on the multi-workgroup kernel: k < k0_q; the
  // trailing block of size n - k0_q is finished by one workgroup
  const int tail = tridiag_tail();
  auto stop_of = [&](int n) { return std::max(0, n - tail); };
  std::vector<DevJob> dj;
  dj.reserve(jobs.size() * 2);
  for (int o : order)
    dj.push_back({jobs[o].a_off, jobs[o].w_off, 0, 0, jobs[o].n, 0, o, stop_of(jobs[o].n)});
  // second copy in caller order for the bisection items
  for (size_t o = 0; o < jobs.size(); ++o)
    dj.push_back({jobs[o].a_off, jobs[o].w_off, 0, 0, jobs[o].n, 0, int(o), stop_of(jobs[o].n)});
  std::vector<TailJob> tj;
  int mmax = 0;
  for (int o : order) {
    const int n = jobs[o].n, k0 = stop_of(n);
    tj.push_back({jobs[o].a_off + int64_t(k0) * n + k0, jobs[o].w_off + k0, n - k0, n, k0 > 0 ? (k0 - 1) & 1 : -1, 0});
    mmax = std::max(mmax, n - k0);
  }
  std::vector<BisItem> items;
  constexpr int kBisNT = 256;
  for (size_t o = 0; o < jobs.size(); ++o)
    for (int j0 = 0; j0 < jobs[o].n; j0 += kBisNT / bis_lanes()) items.push_back({int32_t(jobs.size() + o), j0});
  const size_t jb_bytes = dj.size() * sizeof(DevJob);
  const size_t tj_bytes = tj.size() * sizeof(TailJob);
  DBuf<char> meta(ctx_, jb_bytes + tj_bytes + items.size() * sizeof(BisItem));
  ctx_->upload(meta.get(), dj.data(), jb_bytes);
  ctx_->upload(meta.get() + jb_bytes, tj.data(), tj_bytes);
  ctx_->upload(meta.get() + jb_bytes + tj_bytes, items.data(), items.size() * sizeof(BisItem));
  const DevJob* djd = reinterpret_cast<const DevJob*>(meta.get());
  const TailJob* tjd = reinterpret_cast<const TailJob*>(meta.get() + jb_bytes);
  const BisItem* itd = reinterpret_cast<const BisItem*>(meta.get() + jb_bytes + tj_bytes);
  const int nj = int(jobs.size());
  DBuf<double> vp(ctx_, size_t(4 * std::max<int64_t>(w_n, 1)));
  DBuf<TileItem> tiles;
  std::vector<int> ns(jobs.size()), k0s(jobs.size());
  int kmax0 = 0;
  for (size_t o = 0; o < jobs.size(); ++o) {
    ns[o] = jobs[order[o]].n;
    k0s[o] = stop_of(ns[o]);
    kmax0 = std::max(kmax0, k0s[o]);
  }
  // column steps: one cooperative launch for the jobs whose rows fit one
  // XCD's registers, one launch per column for the rest
  const DevJob* djs = djd;
  std::vector<int> k0step = k0s;
  int kstep = kmax0;
  DBuf<DevJob> dstep;
  DBuf<double> rowbuf;
  if (kmax0 > 0 && tridiag_coop_available(ctx_)) {
    std::vector<CoopJob> cj;
    for (size_t o = 0; o < jobs.size(); ++o) cj.push_back({dj[o].a_off, dj[o].w_off, ns[o], k0s[o]});
    rowbuf.reset(ctx_, size_t(2 * std::max<int64_t>(w_n, 1)));
    const int tc = ctx_->timer.begin(ctx_->stream);
    double cbytes = 0, cflops = 0;
    int64_t csteps = 0;
    const std::vector<int> done = tridiag_columns_coop(ctx_, cj, A, D, E, TAU, vp.get(), rowbuf.get(), w_n);
    for (int o : done) {
      // HBM: the matrix read once, the reflectors and the trailing block
      // written once; arithmetic: per column step the rank-2 update and
      // A22 v, 6 flops per trailing element
      const double n = ns[o], k0 = k0s[o];
      cbytes += 8.0 * (n * n + k0 * (n - 1.0 - 0.5 * k0) + (n - k0) * (n - k0));
      for (int k = 0; k < k0s[o]; ++k) cflops += 6.0 * double(ns[o] - k - 1) * double(ns[o] - k - 1);
      csteps = std::max<int64_t>(csteps, k0s[o]);
      k0step[o] = 0;
    }
    // one "launch" per column step of the batch (the steps run in lock-step
    // inside the cooperative launch)
    ctx_->timer.end(tc, ctx_->stream, "tridiag_coop_fp64", cbytes, cflops, csteps);
    if (!done.empty()) {
      std::vector<DevJob> ds(dj.begin(), dj.begin() + jobs.size());
      for (int o : done) ds[o].pad = 0;
      dstep.reset(ctx_, ds.size());
      dstep.upload(ds.data(), ds.size());
      djs = dstep.get();
      kstep = 0;
      for (size_t o = 0; o < jobs.size(); ++o) kstep = std::max(kstep, k0step[o]);
    }
  }
  const int tk = ctx_->timer.begin(ctx_->stream);
  double bytes = 0;
  int64_t launches = 0;
  if (kstep > 0) {
    static const int rows_override = [] {
      const char* e = std::getenv("TK_TRIDIAG_ROWS");
      return e ? std::atoi(e) : 0;
    }();
    static const int64_t rw2_rows = [] {
      const char* e = std::getenv("TK_TRIDIAG_RW2_ROWS");
      return e ? int64_t(std::atoll(e)) : int64_t(4096);
    }();
    for (int k = 0; k < kstep; ++k) {
      // algorithmic traffic: every step reads and writes the trailing block
      int64_t rows = 0, grid4 = 0, grid8 = 0, grid16 = 0, grid8x2 = 0;
      int trail = 0;  // the largest n - k among the active jobs
      for (size_t o = 0; o < ns.size(); ++o)
        if (k < k0step[o]) {
          const int64_t r = ns[o] - k - 1;
          rows += r;
          bytes += 16.0 * double(r) * double(r);
          grid4 += (r + 3) / 4;
          grid8 += (r + 7) / 8;
          grid16 += (r + 15) / 16;
          grid8x2 += (r + 15) / 16;
          trail = std::max(trail, int(r) + 1);
        }
      if (rows == 0) continue;
      // rows per workgroup: 8 (measured best on the D=2048 sweep against 4
      // and 16; TK_TRIDIAG_ROWS overrides, for tuning)
      int r = 8;
      if (rows_override == 4 || rows_override == 8 || rows_override == 16) r = rows_override;
      const size_t lds = 3 * size_t(nmax_) * sizeof(double);
      // two rows per wave once the step has more rows than one round of
      // single-row workgroups holds (TK_TRIDIAG_RW2_ROWS overrides the
      // threshold, for tuning)
      const int rw = (r == 8 && rows > rw2_rows) ? 2 : 1;
      const int64_t grid = rw == 2 ? grid8x2 : r == 16 ? grid16 : r == 8 ? grid8 : grid4;
      // the smallest register capacity that holds the trailing size
      const int cap = trail <= 512 ? 512 : trail <= 1024 ? 1024 : 2048;
      launch_tdstep(r, rw, cap, dim3(unsigned(grid)), lds, ctx_->stream, djs, nj, k, A, D, E, TAU, vp.get(), w_n,
                    nmax_);
      ++launches;
    }
  }
  ctx_->timer.end(tk, ctx_->stream, "tridiag_column_step", bytes, 0.0, launches);
  // the trailing blocks (whole matrices below the threshold), one workgroup each
  const int tt = ctx_->timer.begin(ctx_->stream);
  double tail_bytes = 0;
  for (const TailJob& t : tj) tail_bytes += 8.0 * double(t.m) * double(t.m) * double(t.m) / 3.0;
  if (tail == kRegTail)
    launch_tridiag_reg(ctx_, tjd, nj, A, D, E, TAU, vp.get(), w_n);
  else
    launch_tridiag_t<512, 2, true>(ctx_, tjd, nj, mmax, A, D, E, TAU, vp.get(), w_n);
  TK_HIP(hipGetLastError());
  ctx_->timer.end(tt, ctx_->stream, "tridiag_tail", tail_bytes, 0.0, 1);
  if (kmax0 > 0) {
    // reflectors of the multi-workgroup steps from columns into rows
    std::vector<TileItem> ti;
    for (size_t o = 0; o < jobs.size(); ++o) {
      const int n = ns[o], nt = (n + 31) / 32, k0 = k0s[o];
      if (k0 == 0) continue;
      for (int a2 = 0; a2 < nt; ++a2)
        for (int b2 = 0; b2 <= a2 && 32 * b2 < k0; ++b2) ti.push_back({int32_t(o), a2, b2, 0});
    }
    if (!ti.empty()) {
      tiles.reset(ctx_, ti.size());
      tiles.upload(ti.data(), ti.size());
      hipLaunchKernelGGL(reflector_rows_kernel, dim3(ti.size()), dim3(256), 0, ctx_->stream, tiles.get(), djd, A);
    }
  }
  TK_HIP(hipGetLastError());
  // the TN slot of a job is indexed by its position in the device job array
  DBuf<double> tn_all(ctx_, dj.size());
  const int tb = ctx_->timer.begin(ctx_->stream);
  {
    const size_t lds = 2 * size_t(nmax_) * sizeof(double);
    const int lanes = bis_lanes();
    if (lanes == 4)
      hipLaunchKernelGGL((bisect_kernel<kBisNT, 4>), dim3(items.size()), dim3(kBisNT), lds, ctx_->stream, itd, djd, D,
                         E, W, tn_all.get(), descending ? 1 : 0);
    else if (lanes == 8)
      hipLaunchKernelGGL((bisect_kernel<kBisNT, 8>), dim3(items.size()), dim3(kBisNT), lds, ctx_->stream, itd, djd, D,
                         E, W, tn_all.get(), descending ? 1 : 0);
    else
      hipLaunchKernelGGL((bisect_kernel<kBisNT, 16>), dim3(items.size()), dim3(kBisNT), lds, ctx_->stream, itd, djd,
                         D, E, W, tn_all.get(), descending ? 1 : 0);
  }
  ctx_->timer.end(tb, ctx_->stream, "bisection", 0.0, 0.0, 1);
  TK_HIP(hipGetLastError());
  TK_HIP(hipMemcpyAsync(tn_.get(), tn_all.get() + jobs.size(), jobs.size() * sizeof(double),
                        hipMemcpyDeviceToDevice, ctx_->stream));
}

void fast_quotient(Context* ctx, const double* a, const double* b, double* q, int64_t n) {
  if (n < 0) throw ArgError("fast_quotient: negative length");
  if (n == 0) return;
  const int64_t grid = (n + 255) / 256;
  if (grid > int64_t(INT32_MAX)) throw ArgError("fast_quotient: length too large");
  hipLaunchKernelGGL(fast_quot_kernel, dim3(unsigned(grid)), dim3(256), 0, ctx->stream, a, b, q, n);
  TK_HIP(hipGetLastError());
  ctx->sync();
}

void SymEigBatch::vectors(const std::vector<int64_t>& keep, double* Z, int64_t z_n) {
  if (keep.size() != jobs_.size()) throw ArgError("eigensolver: one kept count per job");
  if (Z == A_) throw ArgError("eigensolver: the vector output must not alias the reflectors");
  std::vector<DevJob> dj;
  std::vector<VecItem> vitems;
  int64_t z_tot = 0, lu_tot = 0;
  for (size_t o = 0; o < jobs_.size(); ++o) {
    const SymJob& j = jobs_[o];
    const int64_t k = keep[o];
    if (k < 0 || k > j.n) throw ArgError("eigensolver: kept count outside 0..n");
    if (j.a_off + int64_t(j.n) * j.n > z_n) throw std::out_of_range("eigensolver: vector output too small");
    if (k == 0) continue;
    const int idx = int(dj.size());
    dj.push_back({j.a_off, j.w_off, z_tot, lu_tot, j.n, int32_t(k), int32_t(o), 0});
    for (int64_t c = 0; c < k; ++c) vitems.push_back({int32_t(idx), int32_t(c)});
    z_tot += int64_t(j.n) * k;
    lu_tot += kLuFields * int64_t(j.n) * k;
  }
  if (dj.empty()) return;
  const size_t jb_bytes = dj.size() * sizeof(DevJob);
  const size_t vi_bytes = vitems.size() * sizeof(VecItem);
  DBuf<char> meta(ctx_, jb_bytes + vi_bytes);
  ctx_->upload(meta.get(), dj.data(), jb_bytes);
  ctx_->upload(meta.get() + jb_bytes, vitems.data(), vi_bytes);
  const DevJob* djd = reinterpret_cast<const DevJob*>(meta.get());
  const VecItem* vid = reinterpret_cast<const VecItem*>(meta.get() + jb_bytes);
  // per-job Gershgorin bounds indexed by the caller's job number
  DBuf<double> zi(ctx_, size_t(z_tot)), lu(ctx_, size_t(lu_tot));
  const double* D = de_.get();
  const int64_t w_n = int64_t(de_.size() / 3);
  const double* E = D + w_n;
  const double* TAU = E + w_n;
  const int nv = int(vitems.size());
  const int ti = ctx_->timer.begin(ctx_->stream);
  hipLaunchKernelGGL(invit_kernel, dim3((nv + 63) / 64), dim3(64), 0, ctx_->stream, vid, nv, djd, D, E, W_,
                     tn_.get(), lu.get(), zi.get());
  // per (row, vector): the four factor fields written once, read by two
  // back substitutions and (two of them) by one forward solve; the vector
  // written once and read and written by four passes
  ctx_->timer.end(ti, ctx_->stream, "inverse_iteration", 8.0 * double(z_tot) * 23.0, 0.0, 1);
  TK_HIP(hipGetLastError());
  VecWork work;
  for (const DevJob& j : dj) {
    const double n = j.n, k = j.k;
    work.orth_bytes += 8.0 * (2.0 * n * k * k / orth_block() + 4.0 * n * k);
    work.orth_flops += 4.0 * n * k * k;
    work.bt_bytes += 8.0 * (n * n + 2.0 * k * n * n / kWyNb + 2.0 * n * k);
    work.bt_flops += 2.0 * n * n * k;
  }
  const int to = ctx_->timer.begin(ctx_->stream);
  orth_blocked(ctx_, dj, djd, zi.get(), z_tot, lu.get(), w_n);
  ctx_->timer.end(to, ctx_->stream, "orthonormalize", work.orth_bytes, work.orth_flops, 1);
  const int tb = ctx_->timer.begin(ctx_->stream);
  wy_backtransform(ctx_, dj, djd, A_, a_n_, TAU, zi.get(), z_tot, Z, z_n);
  ctx_->timer.end(tb, ctx_->stream, "backtransform", work.bt_bytes, work.bt_flops, 1);
}

void orthonormalize_columns(Context* ctx, const std::vector<OrthJob>& jobs, double* Z, int64_t z_n) {
  std::vector<DevJob> dj;
  int64_t w = 0;
  for (const OrthJob& j : jobs) {
    if (j.n < 1 || j.n > kSymEigMaxN || j.k < 0 || j.k > j.n || j.first < 0 || j.first > j.k)
      throw ArgError("orthonormalize: bad shape");
    if (j.off < 0 || j.off + int64_t(j.n) * j.k > z_n) throw std::out_of_range("orthonormalize: job outside Z");
    if (j.k == j.first) continue;
    dj.push_back({0, w, j.off, 0, j.n, j.k, int32_t(dj.size()), j.first});
    w += j.k;
  }
  if (dj.empty()) return;
  DBuf<DevJob> d(ctx, dj.size());
  d.upload(dj.data(), dj.size());
  orth_blocked(ctx, dj, d.get(), Z, z_n, nullptr, w);
}

void tridiag_lowest(Context* ctx, const double* alpha, const double* beta, int m, const double* beta_next,
                    double* z, double* theta, double* est) {
  if (m < 1 || m > kTriMax) throw ArgError("tridiag_lowest: order outside 1..512");
  hipLaunchKernelGGL(tridiag_lowest_kernel, dim3(1), dim3(64), 0, ctx->stream, alpha, beta, m, beta_next, z, theta,
                     est);
  TK_HIP(hipGetLastError());
}

void lowest_eigpair(Context* ctx, const double* H, int ld, int m, double* c, double* theta) {
  if (m < 1 || m > 64 || ld < m) throw ArgError("lowest_eigpair: dimension outside 1..64");
  hipLaunchKernelGGL(lowest_eig_kernel, dim3(1), dim3(64), 0, ctx->stream, H, ld, m, c, theta);
  TK_HIP(hipGetLastError());
}

}  // namespace tk
