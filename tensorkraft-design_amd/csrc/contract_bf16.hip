// Split-BF16 variant of the grouped block contraction.
//
// Every FP32 operand value v is split as v = hi + lo with hi = bf16(v) and
// lo = bf16(v - hi); the product is formed from three bf16 MFMAs
// (hi*hi + hi*lo + lo*hi) with FP32 accumulation on gfx950's
// v_mfma_f32_16x16x32_bf16.  The dropped lo*lo term bounds the relative error
// of each product near 2^-16; the Krylov solver wins FP64 accuracy back by
// FP64 residual refinement.  Work-list format, tiling, the 16-byte operand
// runs, the XCD item order and the one launch of a wide list's two item
// lists are those of contract.hip (one wavefront per 32x32 output tile, or
// four per 32x128 workgroup item for the wide outputs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kernels.hpp"
#include "tile_util.hpp"

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

// Items wave0, wave0 + nwaves, ... of the list, one per wavefront pass.
template <typename TO, bool TA, bool TB, bool TC>
__device__ __forceinline__ void bf16_wave_items(const GItem* __restrict__ items, int n_items, int wave0, int nwaves,
                                                const GOut* __restrict__ outs, const GTerm* __restrict__ terms,
                                                const GSrc* __restrict__ srcs, const float* __restrict__ A,
                                                const float* __restrict__ B, TO* __restrict__ C, double beta) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
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
        const int kb = k0 + 8 * lk;
        float a0[8], a1[8], b0[8], b1[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) a0[i] = a1[i] = 0.f;
        for (int s = tm.s_begin; s < tm.s_end; ++s) {
          const GSrc sr = srcs[s];
          const float* __restrict__ Ap = A + sr.a_off;
          const float cf = float(sr.coef);
          float x0[8], x1[8];
          if (!TA) {
            run_load<float, 8>(x0, Ap + int64_t(r0) * sr.lda + kb, r0 < o.m, kb, tm.k);
            run_load<float, 8>(x1, Ap + int64_t(r1) * sr.lda + kb, r1 < o.m, kb, tm.k);
          } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
              const int kk = kb + i;
              x0[i] = (kk < tm.k && r0 < o.m) ? Ap[int64_t(kk) * sr.lda + r0] : 0.f;
              x1[i] = (kk < tm.k && r1 < o.m) ? Ap[int64_t(kk) * sr.lda + r1] : 0.f;
            }
          }
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            a0[i] += cf * x0[i];
            a1[i] += cf * x1[i];
          }
        }
        if (TB) {
          run_load<float, 8>(b0, Bp + int64_t(q0) * tm.ldb + kb, q0 < o.n, kb, tm.k);
          run_load<float, 8>(b1, Bp + int64_t(q1) * tm.ldb + kb, q1 < o.n, kb, tm.k);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int kk = kb + i;
            b0[i] = (kk < tm.k && q0 < o.n) ? Bp[int64_t(kk) * tm.ldb + q0] : 0.f;
            b1[i] = (kk < tm.k && q1 < o.n) ? Bp[int64_t(kk) * tm.ldb + q1] : 0.f;
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

template <typename TO, bool TA, bool TB, bool TC>
__global__ __launch_bounds__(256) void contract_bf16x3_kernel(const GItem* __restrict__ items, int n_items,
                                                              const GOut* __restrict__ outs,
                                                              const GTerm* __restrict__ terms,
                                                              const GSrc* __restrict__ srcs,
                                                              const float* __restrict__ A,
                                                              const float* __restrict__ B, TO* __restrict__ C,
                                                              double beta) {
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  // wave-uniform item index: every descriptor is a scalar load
  const int wave0 = __builtin_amdgcn_readfirstlane(xcd_pos(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                                                   (threadIdx.x >> 6));
  bf16_wave_items<TO, TA, TB, TC>(items, n_items, wave0, nwaves, outs, terms, srcs, A, B, C, beta);
}

// Workgroup items (32 x 128, one wavefront per 32 columns): the
// source-summed A chunk (32 rows x 32 inner indices) is formed once per
// workgroup in LDS (double-buffered, one barrier per chunk), as in the FP64 /
// FP32 workgroup kernel; each wavefront splits it and its own B columns.
template <typename TO, bool TA, bool TB, bool TC>
__device__ __forceinline__ void bf16_wide_items(const GItem* __restrict__ items, int n_items, int blk0, int nblk,
                                                const GOut* __restrict__ outs, const GTerm* __restrict__ terms,
                                                const GSrc* __restrict__ srcs, const float* __restrict__ A,
                                                const float* __restrict__ B, TO* __restrict__ C, double beta) {
  constexpr int KC = 32, PER = 32 * KC / 256;
  __shared__ float As[2][32][KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int buf = 0;
  for (int it = blk0; it < n_items; it += nblk) {
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
        const int kb = k0 + 8 * lk;
        if (TB) {
          run_load<float, 8>(b0, Bp + int64_t(q0) * tm.ldb + kb, live && q0 < o.n, kb, tm.k);
          run_load<float, 8>(b1, Bp + int64_t(q1) * tm.ldb + kb, live && q1 < o.n, kb, tm.k);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int kk = kb + i;
            const bool kin = live && kk < tm.k;
            b0[i] = (kin && q0 < o.n) ? Bp[int64_t(kk) * tm.ldb + q0] : 0.f;
            b1[i] = (kin && q1 < o.n) ? Bp[int64_t(kk) * tm.ldb + q1] : 0.f;
          }
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

template <typename TO, bool TA, bool TB, bool TC>
__global__ __launch_bounds__(256) void contract_bf16x3_wide_kernel(const GItem* __restrict__ items, int n_items,
                                                                   const GOut* __restrict__ outs,
                                                                   const GTerm* __restrict__ terms,
                                                                   const GSrc* __restrict__ srcs,
                                                                   const float* __restrict__ A,
                                                                   const float* __restrict__ B, TO* __restrict__ C,
                                                                   double beta) {
  bf16_wide_items<TO, TA, TB, TC>(items, n_items, xcd_pos(blockIdx.x, gridDim.x), gridDim.x, outs, terms, srcs, A, B,
                                  C, beta);
}

// Both item lists of a wide list in one launch, as contract_fused_kernel.
template <typename TO, bool TA, bool TB, bool TC>
__global__ __launch_bounds__(256) void contract_bf16x3_fused_kernel(const GItem* __restrict__ witems, int n_witems,
                                                                    int wblocks, const GItem* __restrict__ items,
                                                                    int n_items, const GOut* __restrict__ outs,
                                                                    const GTerm* __restrict__ terms,
                                                                    const GSrc* __restrict__ srcs,
                                                                    const float* __restrict__ A,
                                                                    const float* __restrict__ B, TO* __restrict__ C,
                                                                    double beta) {
  if (int(blockIdx.x) < wblocks) {
    bf16_wide_items<TO, TA, TB, TC>(witems, n_witems, xcd_pos(blockIdx.x, wblocks), wblocks, outs, terms, srcs, A, B,
                                    C, beta);
  } else {
    const int nb = int(gridDim.x) - wblocks;
    const int wave0 =
        __builtin_amdgcn_readfirstlane(xcd_pos(int(blockIdx.x) - wblocks, nb) * 4 + int(threadIdx.x >> 6));
    bf16_wave_items<TO, TA, TB, TC>(items, n_items, wave0, nb * 4, outs, terms, srcs, A, B, C, beta);
  }
}

template <typename TO>
void launch_bf16(Context* ctx, const DeviceContractList& L, const float* A, const float* B, TO* C) {
  // 32 x 128 workgroup items have a split-BF16 form; 64 x 64 items run as
  // their 32 x 32 expansion on the wavefront kernel
  const bool wide = L.n_witems > 0 && L.wg_kind == kWgWide;
  if (wide && L.n_narrow > 0 && contract_fuse_wide()) {
    const int wblocks = std::min(L.n_witems, ctx->n_cu * 4);
    const int nblocks = std::min((L.n_narrow + 3) / 4, ctx->n_cu * 8);
#define TK_CASE(F)                                                                                         \
  case F:                                                                                                  \
    hipLaunchKernelGGL((contract_bf16x3_fused_kernel<TO, bool(F & kTransA), bool(F & kTransB), bool(F & kTransC)>), \
                       dim3(wblocks + nblocks), dim3(256), 0, ctx->stream, L.witems, L.n_witems, wblocks, L.items, \
                       L.n_narrow, L.outs, L.terms, L.srcs, A, B, C, L.beta);                           \
    break;
    switch (L.flags & 7) {
      TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7)
    }
#undef TK_CASE
    TK_HIP(hipGetLastError());
    return;
  }
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

template <typename TO>
KernelNode bf16_node(const void* f, dim3 grid, const GItem* items, int n, const DeviceContractList& L, const float* A,
                     const float* B, TO* C) {
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

// The launches of launch_bf16, as graph nodes.
template <typename TO>
std::vector<KernelNode> bf16_nodes(Context* ctx, const DeviceContractList& L, const float* A, const float* B, TO* C) {
  std::vector<KernelNode> out;
  const int F = L.flags & 7;
  const bool wide = L.n_witems > 0 && L.wg_kind == kWgWide;
  if (wide && L.n_narrow > 0 && contract_fuse_wide()) {
    const int wblocks = std::min(L.n_witems, ctx->n_cu * 4);
    const int nblocks = std::min((L.n_narrow + 3) / 4, ctx->n_cu * 8);
    const void* f = nullptr;
#define TK_CASE(G) \
  case G: f = reinterpret_cast<const void*>(&contract_bf16x3_fused_kernel<TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC)>); break;
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
  const int n_wave = wide ? L.n_narrow : L.n_items;
  if (n_wave > 0) {
    const void* f = nullptr;
#define TK_CASE(G) \
  case G: f = reinterpret_cast<const void*>(&contract_bf16x3_kernel<TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC)>); break;
    switch (F) { TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7) }
#undef TK_CASE
    out.push_back(bf16_node(f, dim3(std::min((n_wave + 3) / 4, ctx->n_cu * 8)), L.items, n_wave, L, A, B, C));
  }
  if (wide) {
    const void* f = nullptr;
#define TK_CASE(G) \
  case G: f = reinterpret_cast<const void*>(&contract_bf16x3_wide_kernel<TO, bool(G & kTransA), bool(G & kTransB), bool(G & kTransC)>); break;
    switch (F) { TK_CASE(0) TK_CASE(1) TK_CASE(2) TK_CASE(3) TK_CASE(4) TK_CASE(5) TK_CASE(6) TK_CASE(7) }
#undef TK_CASE
    out.push_back(bf16_node(f, dim3(std::min(L.n_witems, ctx->n_cu * 4)), L.witems, L.n_witems, L, A, B, C));
  }
  return out;
}

}  // namespace

std::vector<KernelNode> contract_bf16x3_nodes(Context* ctx, const DeviceContractList& L, const float* A,
                                              const float* B, float* C) {
  return bf16_nodes<float>(ctx, L, A, B, C);
}
std::vector<KernelNode> contract_bf16x3_nodes(Context* ctx, const DeviceContractList& L, const float* A,
                                              const float* B, double* C) {
  return bf16_nodes<double>(ctx, L, A, B, C);
}

void launch_contract_bf16x3(Context* ctx, const DeviceContractList& L, const float* A, const float* B, float* C) {
  launch_bf16<float>(ctx, L, A, B, C);
}
void launch_contract_bf16x3(Context* ctx, const DeviceContractList& L, const float* A, const float* B, double* C) {
  launch_bf16<double>(ctx, L, A, B, C);
}

}  // namespace tk
