// tk_gemm_noalloc: LinAlgBackend::gemm without any device allocation.
//
// tk_gemm packs its work list (one output, one term, one 32 x 32 tile item
// per output tile) into a device buffer it allocates on every call.  This
// variant packs the same list into `scratch`, a file-scope device array of
// kGemmScratchBytes, and launches the same contraction kernel on it, so the
// results are bit-identical to tk_gemm's.  Each GPU holds its own copy of the
// array until the process ends; its address on the current device is looked
// up once per device (hipGetSymbolAddress) and used everywhere the host needs
// it.  Every call on a device reuses that device's array and waits for its
// kernel before returning, so consecutive calls are safe; calls that overlap
// on one GPU (two host threads with contexts on the same device) race on the
// array.  No counterpart in the reference: for callers that cannot allocate
// device memory (code under stream capture, tight serving loops).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/tk_hip.h"
#include "backend.hpp"
#include "kernels.hpp"

namespace tk {

// 4 MiB of work list: up to 262,000 output tiles of 32 x 32 (e.g. a 16384 x
// 16000 output).
constexpr size_t kGemmScratchBytes = size_t(4) << 20;

static __device__ __align__(256) char scratch[kGemmScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on `device` (made current for the lookup).
char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw ArgError("gemm_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    TK_HIP(hipGetDevice(&cur));
    if (cur != device) TK_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) TK_HIP(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      throw HipError(e, std::string("gemm_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

void dense_gemm_noalloc(Context* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t rsA,
                        int64_t csA, const double* B, int64_t rsB, int64_t csB, double beta, double* C, int64_t rsC,
                        int64_t csC) {
  ContractList cl;
  if (!dense_gemm_list(m, n, k, alpha, rsA, csA, rsB, csB, beta, rsC, csC, &cl)) return;
  const size_t need = DeviceContractList::packed_bytes(cl);
  if (need > kGemmScratchBytes)
    throw ArgError("gemm_noalloc: the " + std::to_string(m) + " x " + std::to_string(n) + " output needs a " +
                   std::to_string(need) + "-byte work list, more than the " + std::to_string(kGemmScratchBytes) +
                   "-byte static scratch");
  if (!C || (!cl.terms.empty() && (!A || !B))) throw ArgError("gemm_noalloc: null operand");
  char* base = scratch_on(ctx->device);
  DeviceContractList d;
  d.upload_into(ctx, cl, base, kGemmScratchBytes);
  launch_contract(ctx, d, A, B, C);
  ctx->sync();  // the array is free for the next call; the host staging copy can go
}

}  // namespace tk
