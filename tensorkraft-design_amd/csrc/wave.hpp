// Wavefront reductions of doubles on CDNA without LDS-crossbar permutes:
// within each row of 16 lanes by DPP moves (partners at xor 1, xor 2, then
// the mirrored 8 and 16 lanes), then the four row results read out with
// v_readlane in row order.  Every lane receives the result, and the order of
// operations is the same in every wavefront (results are bit-identical
// wherever the same values are reduced).  All 64 lanes must be active.
#pragma once
#include <hip/hip_runtime.h>

namespace tk {

template <int ctl>
__device__ __forceinline__ double dpp_move(double x) {
  const int2 v = __builtin_bit_cast(int2, x);
  int2 m;
  m.x = __builtin_amdgcn_update_dpp(0, v.x, ctl, 0xF, 0xF, false);
  m.y = __builtin_amdgcn_update_dpp(0, v.y, ctl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, m);
}

__device__ __forceinline__ double lane_value(double x, int l) {
  const int2 v = __builtin_bit_cast(int2, x);
  int2 r;
  r.x = __builtin_amdgcn_readlane(v.x, l);
  r.y = __builtin_amdgcn_readlane(v.y, l);
  return __builtin_bit_cast(double, r);
}

constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror

__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_move<kDppXor1>(x);
  x += dpp_move<kDppXor2>(x);
  x += dpp_move<kDppHalfMirror>(x);
  x += dpp_move<kDppMirror>(x);
  return ((lane_value(x, 0) + lane_value(x, 16)) + lane_value(x, 32)) + lane_value(x, 48);
}

__device__ __forceinline__ double wave_min(double x) {
  x = fmin(x, dpp_move<kDppXor1>(x));
  x = fmin(x, dpp_move<kDppXor2>(x));
  x = fmin(x, dpp_move<kDppHalfMirror>(x));
  x = fmin(x, dpp_move<kDppMirror>(x));
  return fmin(fmin(lane_value(x, 0), lane_value(x, 16)), fmin(lane_value(x, 32), lane_value(x, 48)));
}

__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_move<kDppXor1>(x));
  x = fmax(x, dpp_move<kDppXor2>(x));
  x = fmax(x, dpp_move<kDppHalfMirror>(x));
  x = fmax(x, dpp_move<kDppMirror>(x));
  return fmax(fmax(lane_value(x, 0), lane_value(x, 16)), fmax(lane_value(x, 32), lane_value(x, 48)));
}

}  // namespace tk
