// XCD-local cooperative Householder column steps; the method is described in
// tdcoop.hpp.  Step k of a job (the notation of the one-launch-per-column
// kernel in tdeig.hip): row k of the trailing matrix with the pending update
// of step k-1 gives the reflector v_k, tau_k (T[k+1][k] = beta); every owned
// row i > k takes the update of step k-1, A22 -= v w^T + w v^T with
// w = tau p - tau^2/2 (p.v) v, and contributes p_k[i] = (A22 v_k)_i.
#include "tdcoop.hpp"
#include "wave.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <numeric>

namespace tk {
namespace {

constexpr int kCoopNT = 512;             // threads per workgroup (8 wavefronts)
constexpr int kCoopWaves = kCoopNT / 64;
constexpr int kCtrStride = 32;           // counters 128 bytes apart
constexpr long kSpinMax = 4000000;       // about 2 s of polling per wait
constexpr int kCoopMaxN = 1024;

struct CoopJobDev {
  int64_t a_off, w_off;
  int32_t n, k0, nwg, fbase;  // fbase: the job's first flag line
};
// workgroup slot: job (-1: idle) and the workgroup's rank in the job
struct CoopSlot {
  int32_t job, rank;
};

// Workgroup sum in a fixed order (identical in every workgroup of a job);
// `red` alternates between two halves so consecutive calls need one barrier.
__device__ __forceinline__ double wg_sum(double x, double* red, int& parity) {
  x = wave_sum(x);
  double* r = red + parity * kCoopWaves;
  parity ^= 1;
  if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < kCoopWaves; ++i) s += r[i];
  return s;
}

// 1/q for a normal q: v_rcp_f64 and two Newton steps (error ~1 ulp).
__device__ __forceinline__ double recip(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  return fma(fma(-q, r, 1.0), r, r);
}

// Loads of data published by other workgroups of the same XCD: agent-scope
// relaxed loads miss the CU-local L1 and read the XCD's L2.
__device__ __forceinline__ double ld_l2(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned xcc_id() {
  // HW_REG_XCC_ID (hwreg 20), bits [3:0]
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11));
}

__global__ void xcc_probe_kernel(unsigned* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

// MaxN bounds n; each wavefront owns RW rows (lane l holds columns l + 64 u).
// Row i (1 <= i < n) belongs to rank (i-1) % nwg, local index (i-1) / nwg =
// wavefront + 8 rr.  Per column step: the pivot row and p_{k-1} from the L2,
// two workgroup reductions (p.v, then the norm beyond the subdiagonal), the
// owned rows branch-free (the LDS vectors are zero beyond n; columns <= k are
// dead and masked out of the products), then the arrival flag and the wait
// for every rank's flag (one line per rank, gathered by one wavefront).
template <int MaxN, int RW>
__global__ __launch_bounds__(kCoopNT) __attribute__((amdgpu_waves_per_eu(4))) void tdcoop_kernel(
    const CoopJobDev* __restrict__ jobs, const CoopSlot* __restrict__ slots, double* __restrict__ A,
    double* __restrict__ Dg, double* __restrict__ Eg, double* __restrict__ TAUg, double* __restrict__ VP,
    double* __restrict__ RB, int64_t w_n, unsigned* __restrict__ flags, unsigned* __restrict__ xctr, int n_xcd,
    int per_xcd, int* __restrict__ err) {
  constexpr int kPL = MaxN / 64;         // own-row entries per lane
  constexpr int kPT = MaxN / kCoopNT;    // vector entries per thread
  __shared__ double vbuf[2][MaxN];       // v of the previous and of this step
  __shared__ double wv[MaxN];            // w of the previous step
  __shared__ double red[2 * kCoopWaves];
  __shared__ double piv[2];              // x_k, x_{k+1} of the updated pivot row
  __shared__ int dead;
  __shared__ int slot_id;
  // this workgroup's slot: the next free one of the XCD it runs on (the
  // launch gives every XCD per_xcd workgroups)
  if (threadIdx.x == 0) {
    const int x = int(xcc_id());
    int l = per_xcd;
    if (x < n_xcd)
      l = int(__hip_atomic_fetch_add(xctr + x * kCtrStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (l >= per_xcd) atomicOr(err, 2);
    slot_id = l < per_xcd ? x * per_xcd + l : -1;
  }
  __syncthreads();
  if (slot_id < 0) return;
  const CoopSlot sl = slots[slot_id];
  if (sl.job < 0) return;
  const CoopJobDev jb = jobs[sl.job];
  const int n = jb.n, k0 = jb.k0, nwg = jb.nwg, g = sl.rank;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront-uniform row indices
  double* a = A + jb.a_off;
  double* d = Dg + jb.w_off;
  double* e = Eg + jb.w_off;
  double* tau = TAUg + jb.w_off;
  unsigned* fl = flags + int64_t(jb.fbase) * kCtrStride;  // rank r's flag at fl[r * kCtrStride]
  for (int j = tid; j < MaxN; j += kCoopNT) {
    vbuf[0][j] = 0.0;
    vbuf[1][j] = 0.0;
    wv[j] = 0.0;
  }
  if (tid == 0) dead = 0;
  // own rows into registers
  double x[RW][kPL];
  int row[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int li = wave + kCoopWaves * rr;
    const int i = 1 + g + nwg * li;
    row[rr] = i < n ? i : n;
#pragma unroll
    for (int u = 0; u < kPL; ++u) {
      const int j = lane + 64 * u;
      x[rr][u] = (i < n && j < n) ? a[int64_t(i) * n + j] : 0.0;
    }
  }
  int parity = 0;
  double tau_prev = 0;
  int cur = 0;  // vbuf[cur] receives v_k; vbuf[cur ^ 1] holds v_{k-1}
  __syncthreads();
  for (int k = 0; k < k0; ++k) {
    double* vp = vbuf[cur ^ 1];
    double* vc = vbuf[cur];
    const bool pend = k > 0;
    // (a) this step's operands: row k (published at step k-1, or A's row 0)
    // and p_{k-1}
    double rk[kPT], pp[kPT];
    const double* rsrc = RB + int64_t((k + 1) & 1) * w_n + jb.w_off;
    const double* psrc = VP + int64_t(2 + ((k + 1) & 1)) * w_n + jb.w_off;
    const double pk = pend ? ld_l2(psrc + k) : 0.0;
#pragma unroll
    for (int u = 0; u < kPT; ++u) {
      const int j = k + tid + kCoopNT * u;
      const bool in = j < n;
      rk[u] = !in ? 0.0 : pend ? ld_l2(rsrc + j) : a[j];
      pp[u] = (pend && in) ? ld_l2(psrc + j) : 0.0;
    }
    // (b) w of step k-1 (v_{k-1}[k] = 1, so w_k = tau p_k + alw) and row k
    // with that update
    double alw = 0.0;
    if (pend) {
      double loc = 0;
#pragma unroll
      for (int u = 0; u < kPT; ++u) {
        const int j = k + tid + kCoopNT * u;
        if (j < n) loc += pp[u] * vp[j];
      }
      const double pv = wg_sum(loc, red, parity);
      alw = -0.5 * tau_prev * tau_prev * pv;
    }
    const double wk = tau_prev * pk + alw;
    double ss = 0;
#pragma unroll
    for (int u = 0; u < kPT; ++u) {
      const int j = k + tid + kCoopNT * u;
      if (j < n) {
        double xv = rk[u];
        if (pend) {
          const double vj = vp[j];
          const double wj = tau_prev * pp[u] + alw * vj;
          wv[j] = wj;
          xv -= wj + wk * vj;
        }
        rk[u] = xv;
        if (j == k) piv[0] = xv;
        if (j == k + 1) piv[1] = xv;
        if (j >= k + 2) ss += xv * xv;
      }
    }
    // (c) the reflector; the reduction's barrier also publishes w and the pivots
    const double sigma = wg_sum(ss, red, parity);
    const double alpha = piv[1], dk = piv[0];
    double tk = 0, beta = alpha, scale = 0;
    if (sigma > 0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0 ? -nrm : nrm;
      // |beta| >= |alpha - beta| / 2 > 0: both quotients through the
      // hardware reciprocal and two Newton steps
      tk = 1.0 - alpha * recip(beta);
      scale = recip(alpha - beta);
    }
#pragma unroll
    for (int u = 0; u < kPT; ++u) {
      const int j = k + tid + kCoopNT * u;
      if (j > k && j < n) vc[j] = j == k + 1 ? 1.0 : rk[u] * scale;
    }
    __syncthreads();
    if (g == 0 && tid == 0) {
      d[k] = dk;
      e[k] = beta;
      tau[k] = tk;
    }
    // (d) own rows: the update of step k-1, then (A22 v_k)_i; the owner of
    // row k+1 publishes it for step k+1
    double* pdst = VP + int64_t(2 + (k & 1)) * w_n + jb.w_off;
    double* rdst = RB + int64_t(k & 1) * w_n + jb.w_off;
    double acc[RW], vi[RW], wi[RW];
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int i = row[rr];
      const bool act = pend && i < n && i > k;
      acc[rr] = 0;
      vi[rr] = act ? vp[i] : 0.0;
      wi[rr] = act ? wv[i] : 0.0;
    }
    // column slots wholly at or below k, or at or beyond n, are skipped by a
    // wavefront-uniform branch; inside a live slot every lane computes (the
    // padding beyond n is zero in the rows and in the LDS vectors, columns
    // <= k are dead and only masked out of the product); the next slot's LDS
    // values are read before this slot's update
    double wn = wv[lane], vn = vp[lane], cn = vc[lane];
#pragma unroll
    for (int u = 0; u < kPL; ++u) {
      const int j = lane + 64 * u;
      const double wj = wn, vj = vn, cj = cn;
      if (u + 1 < kPL) {
        wn = wv[j + 64];
        vn = vp[j + 64];
        cn = vc[j + 64];
      }
      if (64 * u + 63 > k && 64 * u < n) {
        const double mj = j > k ? cj : 0.0;
#pragma unroll
        for (int rr = 0; rr < RW; ++rr) {
          const double xv = fma(-vi[rr], wj, fma(-wi[rr], vj, x[rr][u]));
          x[rr][u] = xv;
          acc[rr] = fma(xv, mj, acc[rr]);
        }
      }
    }
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int i = row[rr];
      if (i < n && i > k) {
        const double sv = wave_sum(acc[rr]);
        if (lane == 0) pdst[i] = sv;
        if (i == k + 1) {
#pragma unroll
          for (int u = 0; u < kPL; ++u) {
            const int j = lane + 64 * u;
            if (j > k && j < n) rdst[j] = x[rr][u];
          }
        }
      }
    }
    tau_prev = tk;
    cur ^= 1;
    // (e) arrival: this workgroup's stores are complete at the L2, then its flag
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) fl[g * kCtrStride] = unsigned(k + 1);
    // the reflector into column k (read only after the kernel)
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int i = row[rr];
      if (lane == 0 && i < n && i >= k + 2) a[int64_t(i) * n + k] = vc[i];
    }
    if (wave == 0) {
      const unsigned target = unsigned(k + 1);
      long s = 0;
      while (true) {
        const unsigned f =
            lane < nwg ? __hip_atomic_load(fl + lane * kCtrStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : target;
        if (__all(f >= target)) break;
        if (++s > kSpinMax) {
          if (lane == 0) {
            dead = 1;
            atomicOr(err, 1);
          }
          break;
        }
      }
    }
    __syncthreads();
    if (dead) return;
  }
  // hand-over to the single-workgroup tail: the trailing rows (update k0-2
  // applied) and v_{k0-1}; A22 v_{k0-1} is in the VP pair already
  const double* vlast = vbuf[cur ^ 1];
  double* vdst = VP + int64_t((k0 + 1) & 1) * w_n + jb.w_off;
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int i = row[rr];
    if (i < n && i >= k0) {
#pragma unroll
      for (int u = 0; u < kPL; ++u) {
        const int j = lane + 64 * u;
        if (j >= k0 && j < n) a[int64_t(i) * n + j] = x[rr][u];
      }
      if (lane == 0) vdst[i] = vlast[i];
    }
  }
}

template <int MaxN, int RW>
const void* coop_fn() {
  return reinterpret_cast<const void*>(&tdcoop_kernel<MaxN, RW>);
}

struct DeviceXcd {
  bool checked = false;
  int n_xcd = 0;  // 0: the dispatch order did not match, the path is off
};
std::mutex g_xcd_mu;
std::mutex g_coop_launch_mu;
DeviceXcd g_xcd[64];

// Number of XCDs, and whether a launch gives each the same number of
// workgroups: recorded for a grid of 64 workgroups per XCD, once per device
// (0: uneven, the cooperative path stays off).
int xcd_count(Context* ctx) {
  std::lock_guard<std::mutex> lk(g_xcd_mu);
  DeviceXcd& dx = g_xcd[ctx->device & 63];
  if (dx.checked) return dx.n_xcd;
  dx.checked = true;
  const int nb = 8 * 64;
  unsigned* dev = nullptr;
  TK_HIP(hipMalloc(&dev, nb * sizeof(unsigned)));
  hipLaunchKernelGGL(xcc_probe_kernel, dim3(nb), dim3(64), 0, ctx->stream, dev);
  std::vector<unsigned> h(nb);
  TK_HIP(hipMemcpyAsync(h.data(), dev, nb * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  TK_HIP(hipStreamSynchronize(ctx->stream));
  TK_HIP(hipFree(dev));
  unsigned mx = 0;
  for (unsigned v : h) mx = std::max(mx, v);
  const int nx = int(mx) + 1;
  // the cooperative launch needs every XCD to receive the same number of
  // workgroups (workgroups find their slots by the XCD they run on)
  std::vector<int> cnt(size_t(nx), 0);
  for (unsigned v : h) ++cnt[v];
  bool ok = nx >= 1 && nx <= 16 && ctx->n_cu % nx == 0;
  for (int c : cnt) ok = ok && c == nb / nx;
  dx.n_xcd = ok ? nx : 0;
  return dx.n_xcd;
}

}  // namespace

bool tridiag_coop_available(Context* ctx) {
  // read per call, so one process can compare both paths
  const char* e = std::getenv("TK_TRIDIAG_COOP");
  if (e && std::atoi(e) == 0) return false;
  return xcd_count(ctx) > 0;
}

std::vector<int> tridiag_columns_coop(Context* ctx, const std::vector<CoopJob>& jobs, double* A, double* D,
                                      double* E, double* TAU, double* VP, double* rowbuf, int64_t w_n) {
  std::vector<int> done;
  if (jobs.empty() || !tridiag_coop_available(ctx)) return done;
  const int nx = xcd_count(ctx);
  // matrices up to kCoopMaxN rows (one XCD's registers hold about 1,000
  // rows of 1,024); 64 VGPRs of rows per lane in both variants
  int nmax = 0;
  for (const CoopJob& j : jobs)
    if (j.n <= kCoopMaxN) nmax = std::max(nmax, j.n);
  if (nmax == 0) return done;
  const int maxn = nmax <= 512 ? 512 : 1024, rw = nmax <= 512 ? 4 : 2;
  const void* fn = maxn == 512 ? coop_fn<512, 4>() : coop_fn<1024, 2>();
  int per_cu = 0;
  TK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kCoopNT, 0));
  const int cap = per_cu * (ctx->n_cu / nx);  // co-resident workgroups per XCD
  const int rows_per_wg = kCoopWaves * rw;
  std::vector<int> nwg(jobs.size());
  std::vector<int> order;
  for (size_t q = 0; q < jobs.size(); ++q) {
    const CoopJob& j = jobs[q];
    if (j.k0 <= 0 || j.n < 3 || j.k0 > j.n - 2 || j.n > kCoopMaxN) continue;
    nwg[q] = (j.n - 1 + rows_per_wg - 1) / rows_per_wg;
    if (cap > 0 && nwg[q] <= cap) order.push_back(int(q));
  }
  if (order.empty()) return done;
  // largest first; each job to the XCD with the most free workgroups, a job
  // that fits nowhere waits for the next launch
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return nwg[p] > nwg[q]; });
  std::vector<int> pending = order;
  DBuf<char> meta;
  DBuf<unsigned> ctr;
  DBuf<int> err(ctx, 1);
  TK_HIP(hipMemsetAsync(err.get(), 0, sizeof(int), ctx->stream));
  while (!pending.empty()) {
    std::vector<int> used(nx, 0);
    std::vector<std::vector<std::pair<int, int>>> per_x(nx);  // (job in this launch, rank)
    std::vector<CoopJobDev> dj;
    std::vector<int> rest;
    int nflag = 0;  // flag lines: one per workgroup of every job
    for (int q : pending) {
      int best = 0;
      for (int x = 1; x < nx; ++x)
        if (used[x] < used[best]) best = x;
      if (used[best] + nwg[q] > cap) {
        rest.push_back(q);
        continue;
      }
      const int id = int(dj.size());
      dj.push_back({jobs[q].a_off, jobs[q].w_off, jobs[q].n, jobs[q].k0, nwg[q], nflag});
      nflag += nwg[q];
      for (int r = 0; r < nwg[q]; ++r) per_x[best].push_back({id, r});
      used[best] += nwg[q];
      done.push_back(q);
    }
    const int wx = *std::max_element(used.begin(), used.end());
    const int grid = wx * nx;
    std::vector<CoopSlot> slots(size_t(grid), CoopSlot{-1, 0});
    for (int x = 0; x < nx; ++x)
      for (size_t l = 0; l < per_x[x].size(); ++l)
        slots[size_t(x) * wx + l] = CoopSlot{per_x[x][l].first, per_x[x][l].second};
    const size_t jb = dj.size() * sizeof(CoopJobDev);
    const size_t sb = slots.size() * sizeof(CoopSlot);
    meta.reset(ctx, jb + sb);
    ctx->upload(meta.get(), dj.data(), jb);
    ctx->upload(meta.get() + jb, slots.data(), sb);
    // arrival flags, then one slot counter per XCD
    const size_t nctr = (size_t(nflag) + size_t(nx)) * kCtrStride;
    ctr.reset(ctx, nctr);
    TK_HIP(hipMemsetAsync(ctr.get(), 0, nctr * sizeof(unsigned), ctx->stream));
    const CoopJobDev* jd = reinterpret_cast<const CoopJobDev*>(meta.get());
    const CoopSlot* sd = reinterpret_cast<const CoopSlot*>(meta.get() + jb);
    unsigned* cp = ctr.get();
    unsigned* xp = cp + size_t(nflag) * kCtrStride;
    int nxa = nx, wxa = wx;
    int* ep = err.get();
    void* args[] = {&jd, &sd, &A, &D, &E, &TAU, &VP, &rowbuf, &w_n, &cp, &xp, &nxa, &wxa, &ep};
    {
      // one cooperative launch in flight per process: two partly resident
      // cooperative grids (contexts on other threads) could wait on each other
      std::lock_guard<std::mutex> lk(g_coop_launch_mu);
      TK_HIP(hipLaunchCooperativeKernel(fn, dim3(unsigned(grid)), dim3(kCoopNT), args, 0, ctx->stream));
      // the tables must outlive the launch
      TK_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (rest.size() == pending.size()) break;  // defensive: nothing placed
    pending.swap(rest);
  }
  int herr = 0;
  TK_HIP(hipMemcpyAsync(&herr, err.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TK_HIP(hipStreamSynchronize(ctx->stream));
  if (herr & 2) throw HipError(hipErrorLaunchFailure, "tridiagonalisation: uneven workgroup placement over the XCDs");
  if (herr) throw HipError(hipErrorLaunchFailure, "tridiagonalisation: a cooperative column-step wait timed out");
  return done;
}

}  // namespace tk
