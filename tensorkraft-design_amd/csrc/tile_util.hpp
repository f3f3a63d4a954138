// Device helpers shared by the contraction kernels (contract.hip,
// contract_bf16.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace tk {

// Item position of workgroup b among nb persistent workgroups.  The
// dispatcher deals workgroups to the eight XCDs in turn (b, b + 8, ... share
// one XCD and its L2), and neighbouring items of a list usually belong to
// one output and share its operand rows and columns; so each XCD takes runs
// of kXcdRun consecutive items, the runs dealt round-robin to the XCDs (the
// list's longest-first order is kept at run granularity).  Workgroups past
// the last whole round of runs keep their own index.  A bijection on
// [0, nb); results do not depend on it.
constexpr int kXcdRun = 16;
static __device__ __forceinline__ int xcd_pos(int b, int nb) {
  const int whole = nb / (8 * kXcdRun) * (8 * kXcdRun);
  if (b >= whole) return b;
  const int x = b & 7, j = b >> 3;
  return (j / kXcdRun) * (8 * kXcdRun) + x * kXcdRun + j % kXcdRun;
}

// The U consecutive elements p[0, U) of one lane (an operand whose inner
// index is contiguous in memory): one or two 16-byte loads when the whole run
// lies inside the block (live, base + U <= n), else element by element with
// zeros past n.  Operand rows are only element-aligned; global loads on this
// target do not need 16-byte alignment.
template <typename T, int U>
static __device__ __forceinline__ void run_load(T (&v)[U], const T* __restrict__ p, bool live, int base, int n) {
  if (live && base + U <= n) {
    __builtin_memcpy(v, p, sizeof(v));
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = (live && base + u < n) ? p[u] : T(0);
  }
}

}  // namespace tk
