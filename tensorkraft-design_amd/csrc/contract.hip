// Persistent grouped block-contraction kernel on CDNA4 MFMA.
//
// Each wavefront owns one 32x32 output tile at a time (2x2 MFMA 16x16 tiles)
// and walks the LPT-ordered item list with a grid stride, so one launch covers
// every sector block of a contraction stage whatever its shape.  Operands are
// read straight into MFMA operand registers, a chunk of 16 (FP64) or 32 (FP32)
// inner indices at a time; the A operand of a term is the
// coefficient-weighted sum of its sources, formed in registers (this is where
// the MPO W_i W_j application is fused).  FP64 uses v_mfma_f64_16x16x4f64,
// FP32 v_mfma_f32_16x16x4f32; split-BF16 lives in contract_bf16.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <string>

#include "contract.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "tile_util.hpp"

namespace tk {

void ContractList::finalize() {
  items.clear();
  witems.clear();
  flops = 0;
  std::vector<GItem> expand;  // 32 x 32 tiling of the wide outputs
  for (size_t o = 0; o < outs.size(); ++o) {
    const GOut& g = outs[o];
    int64_t work = 0;
    for (int t = g.t_begin; t < g.t_end; ++t) {
      const GTerm& tm = terms[t];
      int ns = tm.s_end - tm.s_begin;
      work += int64_t(tm.k) * ns;
      flops += 2.0 * g.m * g.n * tm.k + 2.0 * g.m * tm.k * (ns - 1);
    }
    const bool w = (wg_kind == kWgWide && g.n >= kWideMinN) ||
                   (wg_kind == kWgTile64 && g.m >= tile_min && g.n >= tile_min);
    for (int m0 = 0; m0 < g.m; m0 += 32)
      for (int n0 = 0; n0 < g.n; n0 += 32) {
        int64_t c = std::max<int64_t>(1, work * std::min(32, g.m - m0) * std::min(32, g.n - n0) / 64);
        (w ? expand : items).push_back({int32_t(o), m0, n0, int32_t(std::min<int64_t>(c, INT32_MAX))});
      }
    if (w) {
      const int tm = wg_kind == kWgWide ? 32 : 64, tn = wg_kind == kWgWide ? 128 : 64;
      for (int m0 = 0; m0 < g.m; m0 += tm)
        for (int n0 = 0; n0 < g.n; n0 += tn) {
          int64_t c = std::max<int64_t>(1, work * std::min(tm, g.m - m0) * std::min(tn, g.n - n0) / 64);
          witems.push_back({int32_t(o), m0, n0, int32_t(std::min<int64_t>(c, INT32_MAX))});
        }
    }
  }
  auto lpt = [](const GItem& a, const GItem& b) { return a.cost > b.cost; };
  std::stable_sort(items.begin(), items.end(), lpt);
  std::stable_sort(expand.begin(), expand.end(), lpt);
  std::stable_sort(witems.begin(), witems.end(), lpt);
  n_narrow = int(items.size());
  items.insert(items.end(), expand.begin(), expand.end());
}

void ContractList::validate(int64_t a_elems, int64_t b_elems, int64_t c_elems) const {
  const bool ta = flags & kTransA, tb = flags & kTransB, tc = flags & kTransC;
  auto span = [](int64_t off, int64_t rows, int64_t cols, int64_t ld) {
    // last element touched by a rows x cols row-major slice with leading dim ld
    return off + (rows - 1) * ld + cols;
  };
  for (const GItem& it : items) {
    if (it.out < 0 || it.out >= int32_t(outs.size())) throw std::out_of_range("contract: item out of range");
  }
  for (const GItem& it : witems) {
    if (it.out < 0 || it.out >= int32_t(outs.size())) throw std::out_of_range("contract: item out of range");
  }
  for (const GOut& o : outs) {
    if (o.m <= 0 || o.n <= 0) throw std::out_of_range("contract: empty output block");
    int64_t end = tc ? span(o.c_off, o.n, o.m, o.ldc) : span(o.c_off, o.m, o.n, o.ldc);
    if (o.c_off < 0 || end > c_elems || (tc ? o.ldc < o.m : o.ldc < o.n))
      throw std::out_of_range("contract: output slice outside its buffer");
    if (o.t_begin < 0 || o.t_end > int32_t(terms.size()) || o.t_begin > o.t_end)
      throw std::out_of_range("contract: bad term range");
    for (int t = o.t_begin; t < o.t_end; ++t) {
      const GTerm& tm = terms[t];
      if (tm.k <= 0) throw std::out_of_range("contract: empty inner dimension");
      int64_t bend = tb ? span(tm.b_off, o.n, tm.k, tm.ldb) : span(tm.b_off, tm.k, o.n, tm.ldb);
      if (tm.b_off < 0 || bend > b_elems || (tb ? tm.ldb < tm.k : tm.ldb < o.n))
        throw std::out_of_range("contract: B slice outside its buffer");
      if (tm.s_begin < 0 || tm.s_end > int32_t(srcs.size()) || tm.s_begin >= tm.s_end)
        throw std::out_of_range("contract: bad source range");
      for (int s = tm.s_begin; s < tm.s_end; ++s) {
        const GSrc& sr = srcs[s];
        int64_t aend = ta ? span(sr.a_off, tm.k, o.m, sr.lda) : span(sr.a_off, o.m, tm.k, sr.lda);
        if (sr.a_off < 0 || aend > a_elems || (ta ? sr.lda < o.m : sr.lda < tm.k))
          throw std::out_of_range("contract: A slice outside its buffer");
      }
    }
  }
}

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  using acc_t = v4d;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator element r of lane group g holds row g + 4r
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct Mfma<float> {
  using acc_t = v4f;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // accumulator element r of lane group g holds row 4g + r
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// Inner-dimension chunks of 4*U: lane group lk supplies k = k0 + U*lk + u at
// MFMA step u (a permutation of the k order inside the chunk, the same for A
// and B), so each lane reads U consecutive k of a row-major A row.  The work
// item index is made wave-uniform so every descriptor (item, output, term,
// source) is a scalar load, and the sources of a term are taken two at a
// time so that the operand loads of a chunk are all in flight before the
// first FMA.
// Items wave0, wave0 + nwaves, ... of the list, one per wavefront pass.
template <typename TI, typename TO, bool TA, bool TB, bool TC, int U>
__device__ __forceinline__ void wave_items(const GItem* __restrict__ items, int n_items, int wave0, int nwaves,
                                           const GOut* __restrict__ outs, const GTerm* __restrict__ terms,
                                           const GSrc* __restrict__ srcs, const TI* __restrict__ A,
                                           const TI* __restrict__ B, TO* __restrict__ C, double beta) {
  using M = Mfma<TI>;
  using acc_t = typename M::acc_t;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  for (int it = wave0; it < n_items; it += nwaves) {
    const GItem item = items[it];
    const GOut o = outs[item.out];
    const bool two_m = item.m0 + 16 < o.m;  // wave-uniform: second row tile live
    const bool two_n = item.n0 + 16 < o.n;
    acc_t c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    const int r0 = item.m0 + li, r1 = r0 + 16;
    const int q0 = item.n0 + li, q1 = q0 + 16;
    const bool in_r0 = r0 < o.m, in_r1 = r1 < o.m, in_q0 = q0 < o.n, in_q1 = q1 < o.n;
    // one operand element of source sp at inner index kk (0 outside the block)
    auto lda_at = [&](const TI* __restrict__ Ap, int lda, int row, bool rin, int kk, bool kin) -> TI {
      return (kin && rin) ? (TA ? Ap[int64_t(kk) * lda + row] : Ap[int64_t(row) * lda + kk]) : TI(0);
    };
    for (int t = o.t_begin; t < o.t_end; ++t) {
      const GTerm tm = terms[t];
      const TI* __restrict__ Bp = B + tm.b_off;
      for (int k0 = 0; k0 < tm.k; k0 += 4 * U) {
        const int kb = k0 + U * lk;
        TI a0[U], a1[U], b0[U], b1[U];
        if (TB) {
          run_load<TI, U>(b0, Bp + int64_t(q0) * tm.ldb + kb, in_q0, kb, tm.k);
          run_load<TI, U>(b1, Bp + int64_t(q1) * tm.ldb + kb, in_q1, kb, tm.k);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kk = kb + u;
          const bool kin = kk < tm.k;
          if (!TB) {
            b0[u] = (kin && in_q0) ? Bp[int64_t(kk) * tm.ldb + q0] : TI(0);
            b1[u] = (kin && in_q1) ? Bp[int64_t(kk) * tm.ldb + q1] : TI(0);
          }
          a0[u] = 0;
          a1[u] = 0;
        }
        for (int s = tm.s_begin; s < tm.s_end; s += 2) {
          const GSrc sa = srcs[s];
          const bool two = s + 1 < tm.s_end;
          const GSrc sb = srcs[two ? s + 1 : s];
          const TI* __restrict__ Apa = A + sa.a_off;
          const TI* __restrict__ Apb = A + sb.a_off;
          const TI ca = TI(sa.coef), cb = two ? TI(sb.coef) : TI(0);
          TI xa0[U], xa1[U], xb0[U], xb1[U];
          if (!TA) {
            run_load<TI, U>(xa0, Apa + int64_t(r0) * sa.lda + kb, in_r0, kb, tm.k);
            run_load<TI, U>(xa1, Apa + int64_t(r1) * sa.lda + kb, in_r1, kb, tm.k);
            run_load<TI, U>(xb0, Apb + int64_t(r0) * sb.lda + kb, in_r0 && two, kb, tm.k);
            run_load<TI, U>(xb1, Apb + int64_t(r1) * sb.lda + kb, in_r1 && two, kb, tm.k);
          } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int kk = kb + u;
              const bool kin = kk < tm.k;
              xa0[u] = lda_at(Apa, sa.lda, r0, in_r0, kk, kin);
              xa1[u] = lda_at(Apa, sa.lda, r1, in_r1, kk, kin);
              xb0[u] = lda_at(Apb, sb.lda, r0, in_r0 && two, kk, kin);
              xb1[u] = lda_at(Apb, sb.lda, r1, in_r1 && two, kk, kin);
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            a0[u] += ca * xa0[u];
            a1[u] += ca * xa1[u];
            a0[u] += cb * xb0[u];
            a1[u] += cb * xb1[u];
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          c00 = M::run(a0[u], b0[u], c00);
          if (two_n) c01 = M::run(a0[u], b1[u], c01);
          if (two_m) {
            c10 = M::run(a1[u], b0[u], c10);
            if (two_n) c11 = M::run(a1[u], b1[u], c11);
          }
        }
      }
    }
    TO* __restrict__ Cp = C + o.c_off;
    auto put = [&](int row, int col, TI v) {
      if (row < o.m && col < o.n) {
        TO* dst = TC ? &Cp[int64_t(col) * o.ldc + row] : &Cp[int64_t(row) * o.ldc + col];
        *dst = beta == 0.0 ? TO(v) : TO(v) + TO(beta) * *dst;
      }
    };
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = item.m0 + M::row(lk, r);
      const int col = item.n0 + li;
      put(row, col, c00[r]);
      put(row, col + 16, c01[r]);
      put(row + 16, col, c10[r]);
      put(row + 16, col + 16, c11[r]);
    }
  }
}

template <typename TI, typename TO, bool TA, bool TB, bool TC, int U>
__global__ __launch_bounds__(256) void contract_kernel(const GItem* __restrict__ items, int n_items,
                                                       const GOut* __restrict__ outs,
                                                       const GTerm* __restrict__ terms,
                                                       const GSrc* __restrict__ srcs,
                                                       const TI* __restrict__ A, const TI* __restrict__ B,
                                                       TO* __restrict__ C, double beta) {
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  const int wave0 = __builtin_amdgcn_readfirstlane(xcd_pos(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                                                   (threadIdx.x >> 6));
  wave_items<TI, TO, TA, TB, TC, U>(items, n_items, wave0, nwaves, outs, terms, srcs, A, B, C, beta);
}

// Workgroup items: four wavefronts own a 32 x 128 output tile (wavefront w:
// columns n0 + 32 w ..).  Per inner chunk the workgroup forms the
// source-summed A tile (32 rows x 4U inner indices, sources in list order)
// once in LDS, double-buffered so one barrier per chunk suffices, and every
// wavefront multiplies it by its own B columns held in registers: the A
// operand (in H_eff stage 3 the MPO-weighted sum of T1 blocks) is read once
// per four wavefront tiles.  Same k order and source order as the wavefront
// kernel.
// Items blk0, blk0 + nblk, ... of the list, one per workgroup pass (every
// thread of the workgroup calls this).
template <typename TI, typename TO, bool TA, bool TB, bool TC, int U>
__device__ __forceinline__ void wide_items(const GItem* __restrict__ items, int n_items, int blk0, int nblk,
                                           const GOut* __restrict__ outs, const GTerm* __restrict__ terms,
                                           const GSrc* __restrict__ srcs, const TI* __restrict__ A,
                                           const TI* __restrict__ B, TO* __restrict__ C, double beta) {
  using M = Mfma<TI>;
  using acc_t = typename M::acc_t;
  constexpr int KC = 4 * U;              // inner indices per chunk
  constexpr int PER = 32 * KC / 256;     // A-tile elements per thread
  __shared__ TI As[2][32][KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int buf = 0;
  for (int it = blk0; it < n_items; it += nblk) {
    const GItem item = items[it];
    const GOut o = outs[item.out];
    const int nw0 = item.n0 + 32 * wv;  // this wavefront's first column
    const bool live = nw0 < o.n;
    const bool two_m = item.m0 + 16 < o.m;
    const bool two_n = nw0 + 16 < o.n;
    acc_t c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    const int q0 = nw0 + li, q1 = q0 + 16;
    const bool in_q0 = live && q0 < o.n, in_q1 = live && q1 < o.n;
    for (int t = o.t_begin; t < o.t_end; ++t) {
      const GTerm tm = terms[t];
      const TI* __restrict__ Bp = B + tm.b_off;
      for (int k0 = 0; k0 < tm.k; k0 += KC) {
        const int kb = k0 + U * lk;
        TI b0[U], b1[U];
        if (TB) {
          run_load<TI, U>(b0, Bp + int64_t(q0) * tm.ldb + kb, in_q0, kb, tm.k);
          run_load<TI, U>(b1, Bp + int64_t(q1) * tm.ldb + kb, in_q1, kb, tm.k);
        } else {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int kk = kb + u;
            const bool kin = kk < tm.k;
            b0[u] = (kin && in_q0) ? Bp[int64_t(kk) * tm.ldb + q0] : TI(0);
            b1[u] = (kin && in_q1) ? Bp[int64_t(kk) * tm.ldb + q1] : TI(0);
          }
        }
        TI av[PER];
#pragma unroll
        for (int h = 0; h < PER; ++h) av[h] = TI(0);
        for (int s = tm.s_begin; s < tm.s_end; ++s) {
          const GSrc sr = srcs[s];
          const TI* __restrict__ Ap = A + sr.a_off;
          const TI cf = TI(sr.coef);
          TI x[PER];
#pragma unroll
          for (int h = 0; h < PER; ++h) {
            const int e = tid + 256 * h;
            const int row = TA ? (e & 31) : e / KC, kk = TA ? e / 32 : e % KC;
            const int gr = item.m0 + row, gk = k0 + kk;
            x[h] = (gr < o.m && gk < tm.k) ? (TA ? Ap[int64_t(gk) * sr.lda + gr] : Ap[int64_t(gr) * sr.lda + gk])
                                           : TI(0);
          }
#pragma unroll
          for (int h = 0; h < PER; ++h) av[h] += cf * x[h];
        }
#pragma unroll
        for (int h = 0; h < PER; ++h) {
          const int e = tid + 256 * h;
          const int row = TA ? (e & 31) : e / KC, kk = TA ? e / 32 : e % KC;
          As[buf][row][kk] = av[h];
        }
        __syncthreads();
        if (live) {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const TI a0 = As[buf][li][U * lk + u];
            const TI a1 = As[buf][li + 16][U * lk + u];
            c00 = M::run(a0, b0[u], c00);
            if (two_n) c01 = M::run(a0, b1[u], c01);
            if (two_m) {
              c10 = M::run(a1, b0[u], c10);
              if (two_n) c11 = M::run(a1, b1[u], c11);
            }
          }
        }
        buf ^= 1;
      }
    }
    if (live) {
      TO* __restrict__ Cp = C + o.c_off;
      auto put = [&](int row, int col, TI v) {
        if (row < o.m && col < o.n) {
          TO* dst = TC ? &Cp[int64_t(col) * o.ldc + row] : &Cp[int64_t(row) * o.ldc + col];
          *dst = beta == 0.0 ? TO(v) : TO(v) + TO(beta) * *dst;
        }
      };
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = item.m0 + M::row(lk, r);
        const int col = nw0 + li;
        put(row, col, c00[r]);
        put(row, col + 16, c01[r]);
        put(row + 16, col, c10[r]);
        put(row + 16, col + 16, c11[r]);
      }
    }
  }
}


template <typename TI, typename TO, bool TA, bool TB, bool TC, int U>
__global__ __launch_bounds__(256) void contract_wide_kernel(const GItem* __restrict__ items, int n_items,
                                                            const GOut* __restrict__ outs,
                                                            const GTerm* __restrict__ terms,
                                                            const GSrc* __restrict__ srcs,
                                                            const TI* __restrict__ A, const TI* __restrict__ B,
                                                            TO* __restrict__ C, double beta) {
  wide_items<TI, TO, TA, TB, TC, U>(items, n_items, xcd_pos(blockIdx.x, gridDim.x), gridDim.x, outs, terms, srcs, A,
                                    B, C, beta);
}

// Both item lists of a wide list in one launch: workgroups [0, wblocks) run
// the 32 x 128 items, the others the wavefront items of the narrow outputs,
// so the narrow outputs fill the compute units the workgroup items leave
// idle instead of running as a launch of their own before them.  The two
// lists write disjoint outputs; each item runs the same code as in its own
// kernel, so results are those of the two launches.
template <typename TI, typename TO, bool TA, bool TB, bool TC, int U>
__global__ __launch_bounds__(256) void contract_fused_kernel(const GItem* __restrict__ witems, int n_witems,
                                                             int wblocks, const GItem* __restrict__ items,
                                                             int n_items, const GOut* __restrict__ outs,
                                                             const GTerm* __restrict__ terms,
                                                             const GSrc* __restrict__ srcs,
                                                             const TI* __restrict__ A, const TI* __restrict__ B,
                                                             TO* __restrict__ C, double beta) {
  if (int(blockIdx.x) < wblocks) {
    wide_items<TI, TO, TA, TB, TC, U>(witems, n_witems, xcd_pos(blockIdx.x, wblocks), wblocks, outs, terms, srcs, A,
                                      B, C, beta);
  } else {
    const int nb = int(gridDim.x) - wblocks;
    const int nwaves = nb * 4;
    const int wave0 =
        __builtin_amdgcn_readfirstlane(xcd_pos(int(blockIdx.x) - wblocks, nb) * 4 + int(threadIdx.x >> 6));
    wave_items<TI, TO, TA, TB, TC, U>(items, n_items, wave0, nwaves, outs, terms, srcs, A, B, C, beta);
  }
}

// Workgroup items of 64 x 64: four wavefronts in a 2 x 2 arrangement, each
// owning a 32 x 32 quarter (2 x 2 MFMA tiles).  Per inner chunk of KC = 4U
// indices the workgroup stages the source-summed A chunk (64 x KC) and the B
// chunk (KC x 64) in LDS in MFMA fragment order ([16-row group][step][lane],
// so every fragment read is 64 consecutive elements), double-buffered: the
// global loads of chunk c+1 are issued before the MFMAs of chunk c and
// written to the other buffer after them, one barrier per chunk.  Chunks run
// over the terms of the output in list order and over k within a term; the k
// order inside a chunk is the wavefront kernel's (lane group lk supplies
// k = U lk + u at MFMA step u), so A and B see the same permutation.
template <typename TI, typename TO, bool TA, bool TB, bool TC, int U>
__global__ __launch_bounds__(256, 2) void contract_tile64_kernel(const GItem* __restrict__ items, int n_items,
                                                                 const GOut* __restrict__ outs,
                                                                 const GTerm* __restrict__ terms,
                                                                 const GSrc* __restrict__ srcs,
                                                                 const TI* __restrict__ A, const TI* __restrict__ B,
                                                                 TO* __restrict__ C, double beta) {
  using M = Mfma<TI>;
  using acc_t = typename M::acc_t;
  constexpr int KC = 4 * U;
  constexpr int PA = 64 * KC / 256;  // staged elements per thread and chunk, A and B alike
  __shared__ TI As[2][4][U][64];
  __shared__ TI Bs[2][4][U][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;  // this wavefront's 32 x 32 quarter
  const int li = lane & 15, lk = lane >> 4;
  // the (row or column, k) element each thread stages, and its fragment slot
  int a_r[PA], a_k[PA], a_slot[PA], b_c[PA], b_k[PA], b_slot[PA];
#pragma unroll
  for (int h = 0; h < PA; ++h) {
    const int e = tid + 256 * h;
    a_r[h] = TA ? (e & 63) : e / KC;
    a_k[h] = TA ? e / 64 : e % KC;
    b_c[h] = TB ? e / KC : (e & 63);
    b_k[h] = TB ? e % KC : e / 64;
    a_slot[h] = ((a_r[h] >> 4) * U + a_k[h] % U) * 64 + (a_r[h] & 15) + 16 * (a_k[h] / U);
    b_slot[h] = ((b_c[h] >> 4) * U + b_k[h] % U) * 64 + (b_c[h] & 15) + 16 * (b_k[h] / U);
  }
  for (int it = xcd_pos(blockIdx.x, gridDim.x); it < n_items; it += gridDim.x) {
    const GItem item = items[it];
    const GOut o = outs[item.out];
    acc_t c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    // chunk cursor over (term, k0)
    int t = o.t_begin, k0 = 0;
    TI av[PA], bv[PA];
    auto load = [&](int tt, int kk0) {
      const GTerm tm = terms[tt];
      const TI* __restrict__ Bp = B + tm.b_off;
#pragma unroll
      for (int h = 0; h < PA; ++h) {
        const int gc = item.n0 + b_c[h], gk = kk0 + b_k[h];
        bv[h] = (gc < o.n && gk < tm.k) ? (TB ? Bp[int64_t(gc) * tm.ldb + gk] : Bp[int64_t(gk) * tm.ldb + gc])
                                        : TI(0);
        av[h] = TI(0);
      }
      for (int s = tm.s_begin; s < tm.s_end; ++s) {
        const GSrc sr = srcs[s];
        const TI* __restrict__ Ap = A + sr.a_off;
        const TI cf = TI(sr.coef);
#pragma unroll
        for (int h = 0; h < PA; ++h) {
          const int gr = item.m0 + a_r[h], gk = kk0 + a_k[h];
          const TI x = (gr < o.m && gk < tm.k) ? (TA ? Ap[int64_t(gk) * sr.lda + gr] : Ap[int64_t(gr) * sr.lda + gk])
                                               : TI(0);
          av[h] += cf * x;
        }
      }
    };
    auto store = [&](int buf) {
      TI* as = &As[buf][0][0][0];
      TI* bs = &Bs[buf][0][0][0];
#pragma unroll
      for (int h = 0; h < PA; ++h) {
        as[a_slot[h]] = av[h];
        bs[b_slot[h]] = bv[h];
      }
    };
    if (t < o.t_end) {
      load(t, k0);
      store(0);
    }
    __syncthreads();
    int buf = 0;
    while (t < o.t_end) {
      // advance the cursor and request the next chunk before this one's MFMAs
      int tn = t, kn = k0 + KC;
      if (kn >= terms[t].k) {
        ++tn;
        kn = 0;
      }
      const bool more = tn < o.t_end;
      if (more) load(tn, kn);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const TI a0 = As[buf][2 * wr][u][lane], a1 = As[buf][2 * wr + 1][u][lane];
        const TI b0 = Bs[buf][2 * wc][u][lane], b1 = Bs[buf][2 * wc + 1][u][lane];
        c00 = M::run(a0, b0, c00);
        c01 = M::run(a0, b1, c01);
        c10 = M::run(a1, b0, c10);
        c11 = M::run(a1, b1, c11);
      }
      if (more) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
      t = tn;
      k0 = kn;
    }
    TO* __restrict__ Cp = C + o.c_off;
    auto put = [&](int row, int col, TI v) {
      if (row < o.m && col < o.n) {
        TO* dst = TC ? &Cp[int64_t(col) * o.ldc + row] : &Cp[int64_t(row) * o.ldc + col];
        *dst = beta == 0.0 ? TO(v) : TO(v) + TO(beta) * *dst;
      }
    };
    const int rb = item.m0 + 32 * wr, cb = item.n0 + 32 * wc;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rb + M::row(lk, r);
      const int col = cb + li;
      put(row, col, c00[r]);
      put(row, col + 16, c01[r]);
      put(row + 16, col, c10[r]);
      put(row + 16, col + 16, c11[r]);
    }
  }
}

template <typename T> struct KUnroll;
template <> struct KUnroll<double> { static constexpr int value = 4; };
template <> struct KUnroll<float> { static constexpr int value = 8; };


template <typename TI, typename TO>
void launch_typed(Context* ctx, const DeviceContractList& L, const TI* A, const TI* B, TO* C) {
  if (L.n_witems > 0 && L.wg_kind == kWgWide && L.n_narrow > 0 && contract_fuse_wide()) {
    const int wblocks = std::min(L.n_witems, ctx->n_cu * 4);
    const int nblocks = std::min((L.n_narrow + 3) / 4, ctx->n_cu * 8);
#define TK_CASE(F)                                                                                     \
  case F:                                                                                              \
    hipLaunchKernelGGL((contract_fused_kernel<TI, TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC), KUnroll<TI>::value>), \
                       dim3(wblocks + nblocks), dim3(256), 0, ctx->stream, L.witems, L.n_witems, wblocks, L.items, \
                       L.n_narrow, L.outs, L.terms, L.srcs, A, B, C, L.beta);                          \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
    TK_HIP(hipGetLastError());
    return;
  }
  // wavefront items: all of them, or only the narrow ones when the list has
  // workgroup items for its wide outputs
  const int n_wave = L.n_witems > 0 ? L.n_narrow : L.n_items;
  if (n_wave > 0) {
    const int waves_per_block = 4;
    int blocks = (n_wave + waves_per_block - 1) / waves_per_block;
    blocks = std::min(blocks, ctx->n_cu * 8);
    dim3 grid(blocks), block(64 * waves_per_block);
#define TK_CASE(F)                                                                                     \
  case F:                                                                                              \
    hipLaunchKernelGGL((contract_kernel<TI, TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC), KUnroll<TI>::value>), \
                       grid, block, 0, ctx->stream, L.items, n_wave, L.outs, L.terms, L.srcs, A, B, C, L.beta); \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
  }
  if (L.n_witems > 0 && L.wg_kind == kWgWide) {
    const int blocks = std::min(L.n_witems, ctx->n_cu * 4);
#define TK_CASE(F)                                                                                     \
  case F:                                                                                              \
    hipLaunchKernelGGL((contract_wide_kernel<TI, TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC), KUnroll<TI>::value>), \
                       dim3(blocks), dim3(256), 0, ctx->stream, L.witems, L.n_witems, L.outs, L.terms, L.srcs, A, B, C, \
                       L.beta);                                                                        \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
  } else if (L.n_witems > 0 && L.wg_kind == kWgTile64) {
    const int blocks = std::min(L.n_witems, ctx->n_cu * 4);
#define TK_CASE(F)                                                                                     \
  case F:                                                                                              \
    hipLaunchKernelGGL((contract_tile64_kernel<TI, TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC), KUnroll<TI>::value>), \
                       dim3(blocks), dim3(256), 0, ctx->stream, L.witems, L.n_witems, L.outs, L.terms, L.srcs, A, B, C, \
                       L.beta);                                                                        \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
  }
  TK_HIP(hipGetLastError());
}


template <typename TI, typename TO>
KernelNode make_node(const void* f, dim3 grid, const GItem* items, int n, const DeviceContractList& L, const TI* A,
                     const TI* B, TO* C) {
  KernelNode k;
  k.func = f;
  k.grid = grid;
  k.block = dim3(256);
  k.nargs = 9;
  k.set(0, items);
  k.set(1, n);
  k.set(2, L.outs);
  k.set(3, L.terms);
  k.set(4, L.srcs);
  k.set(5, A);
  k.set(6, B);
  k.set(7, C);
  k.set(8, L.beta);
  k.a_arg = 5;
  k.b_arg = 6;
  k.c_arg = 7;
  k.bind();
  return k;
}

// The same kernels, shapes and arguments as launch_typed, as graph nodes.
template <typename TI, typename TO>
std::vector<KernelNode> nodes_typed(Context* ctx, const DeviceContractList& L, const TI* A, const TI* B, TO* C) {
  std::vector<KernelNode> out;
  const int F = L.flags & 7;
  if (L.n_witems > 0 && L.wg_kind == kWgWide && L.n_narrow > 0 && contract_fuse_wide()) {
    const int wblocks = std::min(L.n_witems, ctx->n_cu * 4);
    const int nblocks = std::min((L.n_narrow + 3) / 4, ctx->n_cu * 8);
    const void* f = nullptr;
#define TK_CASE(G) \
  case G: f = reinterpret_cast<const void*>(&contract_fused_kernel<TI, TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC), KUnroll<TI>::value>); break;
    switch (F) { TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7) }
#undef TK_CASE
    KernelNode k;
    k.func = f;
    k.grid = dim3(wblocks + nblocks);
    k.block = dim3(256);
    k.nargs = 12;
    k.set(0, L.witems);
    k.set(1, L.n_witems);
    k.set(2, wblocks);
    k.set(3, L.items);
    k.set(4, L.n_narrow);
    k.set(5, L.outs);
    k.set(6, L.terms);
    k.set(7, L.srcs);
    k.set(8, A);
    k.set(9, B);
    k.set(10, C);
    k.set(11, L.beta);
    k.a_arg = 8;
    k.b_arg = 9;
    k.c_arg = 10;
    k.bind();
    out.push_back(k);
    return out;
  }
  const int n_wave = L.n_witems > 0 ? L.n_narrow : L.n_items;
  if (n_wave > 0) {
    const int blocks = std::min((n_wave + 3) / 4, ctx->n_cu * 8);
    const void* f = nullptr;
#define TK_CASE(G) \
  case G: f = reinterpret_cast<const void*>(&contract_kernel<TI, TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC), KUnroll<TI>::value>); break;
    switch (F) { TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7) }
#undef TK_CASE
    out.push_back(make_node(f, dim3(blocks), L.items, n_wave, L, A, B, C));
  }
  if (L.n_witems > 0) {
    const bool wide = L.wg_kind == kWgWide;
    const int blocks = std::min(L.n_witems, ctx->n_cu * 4);
    const void* f = nullptr;
#define TK_CASE(G)                                                                                                  \
  case G:                                                                                                           \
    f = wide ? reinterpret_cast<const void*>(&contract_wide_kernel<TI, TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC), KUnroll<TI>::value>) \
             : reinterpret_cast<const void*>(&contract_tile64_kernel<TI, TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC), KUnroll<TI>::value>); \
    break;
    switch (F) { TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7) }
#undef TK_CASE
    out.push_back(make_node(f, dim3(blocks), L.witems, L.n_witems, L, A, B, C));
  }
  return out;
}

}  // namespace

namespace {
struct Packing {
  size_t o_items, o_witems, o_outs, o_terms, o_srcs, total;
  explicit Packing(const ContractList& h) {
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    o_items = 0;
    o_witems = align(o_items + h.items.size() * sizeof(GItem));
    o_outs = align(o_witems + h.witems.size() * sizeof(GItem));
    o_terms = align(o_outs + h.outs.size() * sizeof(GOut));
    o_srcs = align(o_terms + h.terms.size() * sizeof(GTerm));
    total = align(o_srcs + h.srcs.size() * sizeof(GSrc));
  }
};
}  // namespace

size_t DeviceContractList::packed_bytes(const ContractList& h) { return Packing(h).total; }

bool contract_fuse_wide() {
  const char* e = std::getenv("TK_CONTRACT_FUSE");
  return !(e && e[0] == '0');
}

void DeviceContractList::pack(const ContractList& h, char* dst, char* base) {
  const Packing pk(h);
  std::copy_n(reinterpret_cast<const char*>(h.items.data()), h.items.size() * sizeof(GItem), dst + pk.o_items);
  std::copy_n(reinterpret_cast<const char*>(h.witems.data()), h.witems.size() * sizeof(GItem), dst + pk.o_witems);
  std::copy_n(reinterpret_cast<const char*>(h.outs.data()), h.outs.size() * sizeof(GOut), dst + pk.o_outs);
  std::copy_n(reinterpret_cast<const char*>(h.terms.data()), h.terms.size() * sizeof(GTerm), dst + pk.o_terms);
  std::copy_n(reinterpret_cast<const char*>(h.srcs.data()), h.srcs.size() * sizeof(GSrc), dst + pk.o_srcs);
  items = reinterpret_cast<GItem*>(base + pk.o_items);
  witems = reinterpret_cast<GItem*>(base + pk.o_witems);
  outs = reinterpret_cast<GOut*>(base + pk.o_outs);
  terms = reinterpret_cast<GTerm*>(base + pk.o_terms);
  srcs = reinterpret_cast<GSrc*>(base + pk.o_srcs);
  n_items = int(h.items.size());
  n_narrow = h.witems.empty() ? n_items : h.n_narrow;
  n_witems = int(h.witems.size());
  wg_kind = h.witems.empty() ? kWgNone : h.wg_kind;
  flags = h.flags;
  flops = h.flops;
  beta = h.beta;
}

void DeviceContractList::upload_into(Context* ctx, const ContractList& h, char* base, size_t capacity) {
  const Packing pk(h);
  if (pk.total > capacity)
    throw ArgError("contraction list of " + std::to_string(pk.total) + " bytes exceeds the " +
                   std::to_string(capacity) + "-byte buffer");
  host.assign(pk.total, 0);
  pack(h, host.data(), base);
  ctx->upload(base, host.data(), pk.total);
}

int ListPack::add(ContractList&& l, int64_t a_n, int64_t b_n, int64_t c_n) {
  if (l.outs.empty()) return -1;
  l.finalize();
  l.validate(a_n, b_n, c_n);
  lists_.push_back(std::move(l));
  return int(lists_.size()) - 1;
}

void ListPack::upload(Context* ctx) {
  dev_.clear();
  dev_.resize(lists_.size());
  std::vector<size_t> off(lists_.size() + 1, 0);
  for (size_t i = 0; i < lists_.size(); ++i) off[i + 1] = off[i] + DeviceContractList::packed_bytes(lists_[i]);
  if (off.back() == 0) return;
  buf_.reset(ctx, off.back());
  host_.assign(off.back(), 0);
  for (size_t i = 0; i < lists_.size(); ++i) dev_[i].pack(lists_[i], host_.data() + off[i], buf_.get() + off[i]);
  ctx->upload(buf_.get(), host_.data(), off.back());
}

void DeviceContractList::upload(Context* ctx, const ContractList& h) {
  const size_t total = Packing(h).total;
  buf.reset(ctx, total);
  upload_into(ctx, h, buf.get(), total);
}

void launch_contract(Context* ctx, const DeviceContractList& L, const double* A, const double* B, double* C) {
  launch_typed<double, double>(ctx, L, A, B, C);
}
void launch_contract(Context* ctx, const DeviceContractList& L, const float* A, const float* B, float* C) {
  launch_typed<float, float>(ctx, L, A, B, C);
}
void launch_contract(Context* ctx, const DeviceContractList& L, const float* A, const float* B, double* C) {
  launch_typed<float, double>(ctx, L, A, B, C);
}
std::vector<KernelNode> contract_nodes(Context* ctx, const DeviceContractList& L, const double* A, const double* B,
                                       double* C) {
  return nodes_typed<double, double>(ctx, L, A, B, C);
}
std::vector<KernelNode> contract_nodes(Context* ctx, const DeviceContractList& L, const float* A, const float* B,
                                       double* C) {
  return nodes_typed<float, double>(ctx, L, A, B, C);
}
std::vector<KernelNode> contract_nodes(Context* ctx, const DeviceContractList& L, const float* A, const float* B,
                                       float* C) {
  return nodes_typed<float, float>(ctx, L, A, B, C);
}

}  // namespace tk
