// Host-callable launchers for the HIP kernels of the library.
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

#include "contract.hpp"
#include "device.hpp"

namespace tk {

// A ContractList resident on the device (one packed allocation).
struct DeviceContractList {
  DBuf<char> buf;
  std::vector<char> host;  // kept alive until the list is released
  const GItem* items = nullptr;
  const GOut* outs = nullptr;
  const GTerm* terms = nullptr;
  const GSrc* srcs = nullptr;
  const GItem* witems = nullptr;
  int n_items = 0;
  int n_narrow = 0;
  int n_witems = 0;
  int wg_kind = kWgNone;  // what the workgroup items are (WgKind)
  int flags = 0;
  double flops = 0;
  double beta = 0;  // C = sum + beta * C
  void upload(Context* ctx, const ContractList& h);
  // The same packing into caller-owned device memory of `capacity` bytes
  // (no allocation); throws ArgError when the list does not fit.
  void upload_into(Context* ctx, const ContractList& h, char* dst, size_t capacity);
  // bytes the packed list of h takes
  static size_t packed_bytes(const ContractList& h);
  // The packed bytes of h into host memory at dst, and the table pointers
  // set for the same bytes at device address base (no transfer).
  void pack(const ContractList& h, char* dst, char* base);
};

// Several lists in one device allocation, uploaded with one copy, for
// phases that launch many small contractions in sequence.
class ListPack {
 public:
  // Finalizes `l` and checks it against the element counts of the buffers it
  // will run on; its index in the pack (-1 for a list without outputs, which
  // launches nothing).
  int add(ContractList&& l, int64_t a_n, int64_t b_n, int64_t c_n);
  // One allocation and one host-to-device copy for every list added.
  void upload(Context* ctx);
  const DeviceContractList& operator[](int i) const { return dev_[size_t(i)]; }

 private:
  std::vector<ContractList> lists_;
  std::vector<DeviceContractList> dev_;
  std::vector<char> host_;
  DBuf<char> buf_;
};

// One kernel launch held as a graph node: the function, its shape and its
// argument values, each in an 8-byte slot that args[i] points to (an int
// occupies the slot's low bytes).  a_arg, b_arg and c_arg name the slots of
// a contraction's A, B and C operands (-1 for other kernels).
struct KernelNode {
  const void* func = nullptr;
  dim3 grid, block;
  unsigned shmem = 0;
  int nargs = 0;
  int a_arg = -1, b_arg = -1, c_arg = -1;
  uint64_t slot[12] = {};
  void* args[12] = {};
  template <typename T>
  void set(int i, T v) {
    static_assert(sizeof(T) <= sizeof(uint64_t), "argument slot");
    slot[i] = 0;
    std::memcpy(&slot[i], &v, sizeof(T));
  }
  void bind() {
    for (int i = 0; i < nargs; ++i) args[i] = &slot[i];
  }
};

// Whether a wide list runs its workgroup and wavefront items as one launch
// (TK_CONTRACT_FUSE, read per launch or graph build; default on, 0: two).
bool contract_fuse_wide();
void launch_contract(Context* ctx, const DeviceContractList& L, const double* A, const double* B, double* C);
// The launches launch_contract would make, as graph nodes (their a_arg,
// b_arg and c_arg name the A, B and C slots).
std::vector<KernelNode> contract_nodes(Context* ctx, const DeviceContractList& L, const double* A, const double* B,
                                       double* C);
std::vector<KernelNode> contract_nodes(Context* ctx, const DeviceContractList& L, const float* A, const float* B,
                                       double* C);
std::vector<KernelNode> contract_nodes(Context* ctx, const DeviceContractList& L, const float* A, const float* B,
                                       float* C);
// convert_f64_to_f32 as a graph node (argument 0 is x)
KernelNode convert_f64_to_f32_node(Context* ctx, const double* x, float* y, int64_t n);
void launch_contract(Context* ctx, const DeviceContractList& L, const float* A, const float* B, float* C);
void launch_contract(Context* ctx, const DeviceContractList& L, const float* A, const float* B, double* C);
// split-BF16 (hi+lo, three products, FP32 accumulation) on FP32 operands
void launch_contract_bf16x3(Context* ctx, const DeviceContractList& L, const float* A, const float* B, float* C);
// The launches of launch_contract_bf16x3, as graph nodes (a_arg, b_arg and
// c_arg set as for contract_nodes).
std::vector<KernelNode> contract_bf16x3_nodes(Context* ctx, const DeviceContractList& L, const float* A,
                                              const float* B, float* C);
std::vector<KernelNode> contract_bf16x3_nodes(Context* ctx, const DeviceContractList& L, const float* A,
                                              const float* B, double* C);
void launch_contract_bf16x3(Context* ctx, const DeviceContractList& L, const float* A, const float* B, double* C);

// 2-D strided copies (optionally transposed) in one launch (dmrg.hip).
// dst is m x n with leading dimension dst_ld; src is m x n (trans = 0) or
// n x m (trans = 1, dst(i, j) = src(j, i)) with leading dimension src_ld.
struct CopyJob {
  int64_t src, dst;
  int32_t m, n, src_ld, dst_ld;
  int32_t trans, pad;
};
void launch_copies(Context* ctx, const std::vector<CopyJob>& jobs, const double* src, int64_t src_n, double* dst,
                   int64_t dst_n);
void launch_fill(Context* ctx, double* p, double v, int64_t n);

// ---- vector kernels (vecops.hip) ----
void convert_f64_to_f32(Context* ctx, const double* x, float* y, int64_t n);
void convert_f32_to_f64(Context* ctx, const float* x, double* y, int64_t n);
void fill_f64(Context* ctx, double* x, double v, int64_t n);
// out[i] = sum_j V[i*ldv + j] * x[j] for i < nvec (deterministic two-pass
// reduction); skipped when `active` (a device flag) reads 0
void dot_many(Context* ctx, const double* V, int64_t ldv, int nvec, const double* x, int64_t n, double* out,
              const int* active = nullptr);
// y[j] = beta*y[j] + sum_i c[i] * V[i*ldv + j]; c on the device
void combine(Context* ctx, const double* V, int64_t ldv, int nvec, const double* c, double beta, double* y, int64_t n);
// y[j] -= sum_i h[i] * V[i*ldv + j]; h on the device; skipped when `active` reads 0
void subtract_projection(Context* ctx, const double* V, int64_t ldv, int nvec, const double* h, double* y, int64_t n,
                         const int* active = nullptr);
// t = r or r / (theta - diag) (|theta - diag| > 1e-14) for the Ritz residual
// r = AV c - theta V c of nvec basis vectors, and *rn2 = |r|^2, in one pass
// (c, theta, rn2 on the device; diag may be null)
void ritz_residual(Context* ctx, const double* V, const double* AV, int64_t ldv, int nvec, const double* c,
                   const double* theta, const double* diag, double* t, int64_t n, double* rn2);
// *flag = 1 when a second Gram-Schmidt pass is needed: h[0..nvec) = V^T s and
// h[nvec] = s.s from before the first pass, s.s - |h|^2 < s.s / 2
// The largest basis of the two fused passes below.
int fused_vec_max();
// ritz_residual for nvec <= fused_vec_max() that also gives h[0..nvec) =
// V^T t and h[nvec] = t.t of the t it writes.
void ritz_residual_dots(Context* ctx, const double* V, const double* AV, int64_t ldv, int nvec, const double* c,
                        const double* theta, const double* diag, double* t, int64_t n, double* rn2, double* h);
// y <- y - V h for nvec <= fused_vec_max(), with h_new = V^T y (unless null)
// and *nrm2 = y.y of the result; nothing when *active reads 0.
void project_out_dots(Context* ctx, const double* V, int64_t ldv, int nvec, const double* h, double* y, int64_t n,
                      double* h_new, double* nrm2, const int* active = nullptr);
void reorth_flag(Context* ctx, const double* h, int nvec, int* flag);
// x *= 1/sqrt(*nrm2) read on the device (no-op guard when the norm is zero)
void scale_by_inv_sqrt(Context* ctx, double* x, const double* nrm2, int64_t n);
void scale(Context* ctx, double* x, double a, int64_t n);
void axpy(Context* ctx, double a, const double* x, double* y, int64_t n);
// r = ay - theta*y (all device), optionally t = r / (theta - diag) where |theta-diag| > 1e-14
void residual(Context* ctx, const double* ay, const double* y, double theta, double* r, int64_t n);
void precondition(Context* ctx, const double* r, const double* diag, double theta, double* t, int64_t n);

}  // namespace tk
