// Split-BF16 variant of the grouped block contraction.
//
// Every FP32 operand value v is split as v = hi + lo with hi = bf16(v) and
// lo = bf16(v - hi); the product is formed from three bf16 MFMAs
// (hi*hi + hi*lo + lo*hi) with FP32 accumulation on gfx950's
// v_mfma_f32_16x16x32_bf16.  The dropped lo*lo term bounds the relative error
// of each product near 2^-16; the Krylov solver wins FP64 accuracy back by
// FP64 residual refinement.  Work-list format and tiling are those of
// contract.hip (one wavefront per 32x32 output tile, or four per 32x128
// workgroup item for the wide outputs).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace tk {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh, const bf16x8& bl,
                                     v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

__device__ __forceinline__ void split8(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    __bf16 h = __bf16(v[i]);
    hi[i] = h;
    lo[i] = __bf16(v[i] - float(h));
  }
}

template <typename TO, bool TA, bool TB, bool TC>
__global__ __launch_bounds__(256) void contract_bf16x3_kernel(const GItem* __restrict__ items, int n_items,
                                                              const GOut* __restrict__ outs,
                                                              const GTerm* __restrict__ terms,
                                                              const GSrc* __restrict__ srcs,
                                                              const float* __restrict__ A,
                                                              const float* __restrict__ B, TO* __restrict__ C,
                                                              double beta) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  // wave-uniform item index: every descriptor is a scalar load
  const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  for (int it = wave0; it < n_items; it += nwaves) {
    const GItem item = items[it];
    const GOut o = outs[item.out];
    const bool two_m = item.m0 + 16 < o.m;
    const bool two_n = item.n0 + 16 < o.n;
    v4f c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    const int r0 = item.m0 + li, r1 = r0 + 16;
    const int q0 = item.n0 + li, q1 = q0 + 16;
    for (int t = o.t_begin; t < o.t_end; ++t) {
      const GTerm tm = terms[t];
      const float* __restrict__ Bp = B + tm.b_off;
      for (int k0 = 0; k0 < tm.k; k0 += 32) {
        float a0[8], a1[8], b0[8], b1[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) a0[i] = a1[i] = b0[i] = b1[i] = 0.f;
        for (int s = tm.s_begin; s < tm.s_end; ++s) {
          const GSrc sr = srcs[s];
          const float* __restrict__ Ap = A + sr.a_off;
          const float cf = float(sr.coef);
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int kk = k0 + 8 * lk + i;
            if (kk < tm.k) {
              if (r0 < o.m) a0[i] += cf * (TA ? Ap[int64_t(kk) * sr.lda + r0] : Ap[int64_t(r0) * sr.lda + kk]);
              if (r1 < o.m) a1[i] += cf * (TA ? Ap[int64_t(kk) * sr.lda + r1] : Ap[int64_t(r1) * sr.lda + kk]);
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int kk = k0 + 8 * lk + i;
          if (kk < tm.k) {
            if (q0 < o.n) b0[i] = TB ? Bp[int64_t(q0) * tm.ldb + kk] : Bp[int64_t(kk) * tm.ldb + q0];
            if (q1 < o.n) b1[i] = TB ? Bp[int64_t(q1) * tm.ldb + kk] : Bp[int64_t(kk) * tm.ldb + q1];
          }
        }
        bf16x8 a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l;
        split8(a0, a0h, a0l);
        split8(b0, b0h, b0l);
        c00 = mfma3(a0h, a0l, b0h, b0l, c00);
        if (two_n) {
          split8(b1, b1h, b1l);
          c01 = mfma3(a0h, a0l, b1h, b1l, c01);
        }
        if (two_m) {
          split8(a1, a1h, a1l);
          c10 = mfma3(a1h, a1l, b0h, b0l, c10);
          if (two_n) c11 = mfma3(a1h, a1l, b1h, b1l, c11);
        }
      }
    }
    TO* __restrict__ Cp = C + o.c_off;
    auto put = [&](int row, int col, float v) {
      if (row < o.m && col < o.n) {
        TO* dst = TC ? &Cp[int64_t(col) * o.ldc + row] : &Cp[int64_t(row) * o.ldc + col];
        *dst = beta == 0.0 ? TO(v) : TO(v) + TO(beta) * *dst;
      }
    };
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = item.m0 + 4 * lk + r;
      const int col = item.n0 + li;
      put(row, col, c00[r]);
      put(row, col + 16, c01[r]);
      put(row + 16, col, c10[r]);
      put(row + 16, col + 16, c11[r]);
    }
  }
}

// Workgroup items (32 x 128, one wavefront per 32 columns): the
// source-summed A chunk (32 rows x 32 inner indices) is formed once per
// workgroup in LDS (double-buffered, one barrier per chunk), as in the FP64 /
// FP32 workgroup kernel; each wavefront splits it and its own B columns.
template <typename TO, bool TA, bool TB, bool TC>
__global__ __launch_bounds__(256) void contract_bf16x3_wide_kernel(const GItem* __restrict__ items, int n_items,
                                                                   const GOut* __restrict__ outs,
                                                                   const GTerm* __restrict__ terms,
                                                                   const GSrc* __restrict__ srcs,
                                                                   const float* __restrict__ A,
                                                                   const float* __restrict__ B, TO* __restrict__ C,
                                                                   double beta) {
  constexpr int KC = 32, PER = 32 * KC / 256;
  __shared__ float As[2][32][KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int buf = 0;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const GItem item = items[it];
    const GOut o = outs[item.out];
    const int nw0 = item.n0 + 32 * wv;
    const bool live = nw0 < o.n;
    const bool two_m = item.m0 + 16 < o.m;
    const bool two_n = nw0 + 16 < o.n;
    v4f c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    const int q0 = nw0 + li, q1 = q0 + 16;
    for (int t = o.t_begin; t < o.t_end; ++t) {
      const GTerm tm = terms[t];
      const float* __restrict__ Bp = B + tm.b_off;
      for (int k0 = 0; k0 < tm.k; k0 += KC) {
        float b0[8], b1[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int kk = k0 + 8 * lk + i;
          const bool kin = live && kk < tm.k;
          b0[i] = (kin && q0 < o.n) ? (TB ? Bp[int64_t(q0) * tm.ldb + kk] : Bp[int64_t(kk) * tm.ldb + q0]) : 0.f;
          b1[i] = (kin && q1 < o.n) ? (TB ? Bp[int64_t(q1) * tm.ldb + kk] : Bp[int64_t(kk) * tm.ldb + q1]) : 0.f;
        }
        float av[PER];
#pragma unroll
        for (int h = 0; h < PER; ++h) av[h] = 0.f;
        for (int s = tm.s_begin; s < tm.s_end; ++s) {
          const GSrc sr = srcs[s];
          const float* __restrict__ Ap = A + sr.a_off;
          const float cf = float(sr.coef);
          float x[PER];
#pragma unroll
          for (int h = 0; h < PER; ++h) {
            const int e = tid + 256 * h;
            const int row = TA ? (e & 31) : e / KC, kk = TA ? e / 32 : e % KC;
            const int gr = item.m0 + row, gk = k0 + kk;
            x[h] = (gr < o.m && gk < tm.k) ? (TA ? Ap[int64_t(gk) * sr.lda + gr] : Ap[int64_t(gr) * sr.lda + gk])
                                           : 0.f;
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
          float a0[8], a1[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            a0[i] = As[buf][li][8 * lk + i];
            a1[i] = As[buf][li + 16][8 * lk + i];
          }
          bf16x8 a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l;
          split8(a0, a0h, a0l);
          split8(b0, b0h, b0l);
          c00 = mfma3(a0h, a0l, b0h, b0l, c00);
          if (two_n) {
            split8(b1, b1h, b1l);
            c01 = mfma3(a0h, a0l, b1h, b1l, c01);
          }
          if (two_m) {
            split8(a1, a1h, a1l);
            c10 = mfma3(a1h, a1l, b0h, b0l, c10);
            if (two_n) c11 = mfma3(a1h, a1l, b1h, b1l, c11);
          }
        }
        buf ^= 1;
      }
    }
    if (live) {
      TO* __restrict__ Cp = C + o.c_off;
      auto put = [&](int row, int col, float v) {
        if (row < o.m && col < o.n) {
          TO* dst = TC ? &Cp[int64_t(col) * o.ldc + row] : &Cp[int64_t(row) * o.ldc + col];
          *dst = beta == 0.0 ? TO(v) : TO(v) + TO(beta) * *dst;
        }
      };
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = item.m0 + 4 * lk + r;
        const int col = nw0 + li;
        put(row, col, c00[r]);
        put(row, col + 16, c01[r]);
        put(row + 16, col, c10[r]);
        put(row + 16, col + 16, c11[r]);
      }
    }
  }
}

template <typename TO>
void launch_bf16(Context* ctx, const DeviceContractList& L, const float* A, const float* B, TO* C) {
  // 32 x 128 workgroup items have a split-BF16 form; 64 x 64 items run as
  // their 32 x 32 expansion on the wavefront kernel
  const bool wide = L.n_witems > 0 && L.wg_kind == kWgWide;
  const int n_wave = wide ? L.n_narrow : L.n_items;
  if (n_wave > 0) {
    int blocks = std::min((n_wave + 3) / 4, ctx->n_cu * 8);
    dim3 grid(blocks), block(256);
#define TK_CASE(F)                                                                                         \
  case F:                                                                                                  \
    hipLaunchKernelGGL((contract_bf16x3_kernel<TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC)>), \
                       grid, block, 0, ctx->stream, L.items, n_wave, L.outs, L.terms, L.srcs, A, B, C, L.beta); \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
  }
  if (wide) {
    const int blocks = std::min(L.n_witems, ctx->n_cu * 4);
#define TK_CASE(F)                                                                                         \
  case F:                                                                                                  \
    hipLaunchKernelGGL((contract_bf16x3_wide_kernel<TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC)>), \
                       dim3(blocks), dim3(256), 0, ctx->stream, L.witems, L.n_witems, L.outs, L.terms, L.srcs, A, B, \
                       C, L.beta);                                                                         \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
  }
  TK_HIP(hipGetLastError());
}

}  // namespace

void launch_contract_bf16x3(Context* ctx, const DeviceContractList& L, const float* A, const float* B, float* C) {
  launch_bf16<float>(ctx, L, A, B, C);
}
void launch_contract_bf16x3(Context* ctx, const DeviceContractList& L, const float* A, const float* B, double* C) {
  launch_bf16<double>(ctx, L, A, B, C);
}

}  // namespace tk
