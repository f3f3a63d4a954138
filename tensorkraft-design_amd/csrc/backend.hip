// Dense and rank-2 block-sparse backend operations of the LinAlgBackend /
// SparseLinAlgBackend traits (tk-linalg/src/traits.rs:25-205) on the
// grouped-contraction kernel and the batched Jacobi eigensolver.
#include "backend.hpp"

#include <string>

#include "truncate.hpp"

#include <algorithm>
#include <cmath>
#include <map>

#include "tdeig.hpp"
#include "kernels.hpp"

namespace tk {
namespace {

// Describe a strided operand as a row-major slice (ld) or its transpose.
bool as_slice(int64_t rows, int64_t cols, int64_t rs, int64_t cs, int64_t* ld, bool* trans) {
  if (cs == 1 || cols == 1) {
    *ld = rows == 1 ? std::max<int64_t>(cols, 1) : rs;
    *trans = false;
    return *ld >= cols || rows == 1;
  }
  if (rs == 1 || rows == 1) {
    *ld = cols == 1 ? std::max<int64_t>(rows, 1) : cs;
    *trans = true;
    return *ld >= rows || cols == 1;
  }
  return false;
}

void run(Context* ctx, ContractList& cl, int64_t a_n, int64_t b_n, int64_t c_n, const double* A, const double* B,
         double* C) {
  cl.finalize();
  cl.validate(a_n, b_n, c_n);
  DeviceContractList d;
  d.upload(ctx, cl);
  launch_contract(ctx, d, A, B, C);
  ctx->sync();
}

}  // namespace

bool dense_gemm_list(int64_t m, int64_t n, int64_t k, double alpha, int64_t rsA, int64_t csA, int64_t rsB,
                     int64_t csB, double beta, int64_t rsC, int64_t csC, ContractList* cl) {
  if (m < 0 || n < 0 || k < 0) throw ArgError("gemm: negative dimension");
  if (m == 0 || n == 0) return false;
  if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) throw ArgError("gemm: dimension above 2^31 - 1");
  int64_t lda, ldb, ldc;
  bool ta, tb, tc;
  if (!as_slice(m, k, rsA, csA, &lda, &ta) || !as_slice(k, n, rsB, csB, &ldb, &tb) ||
      !as_slice(m, n, rsC, csC, &ldc, &tc))
    throw ArgError("gemm: operand without a unit stride (NonContiguousForBlas)");
  cl->clear();
  cl->flags = (ta ? kTransA : 0) | (tb ? kTransB : 0) | (tc ? kTransC : 0);
  cl->beta = beta;
  cl->begin_out(0, int(m), int(n), int(ldc));
  // C = beta C when there is nothing to add: an output with no terms scales (beta = 0 zeroes)
  if (k > 0 && alpha != 0.0) {
    cl->begin_term(0, int(ldb), int(k));
    cl->add_src(0, int(lda), alpha);
  }
  cl->finalize();
  const int64_t c_n = tc ? (n - 1) * ldc + m : (m - 1) * ldc + n;
  int64_t a_n = 0, b_n = 0;
  if (k > 0 && alpha != 0.0) {
    a_n = ta ? (k - 1) * lda + m : (m - 1) * lda + k;
    b_n = tb ? (n - 1) * ldb + k : (k - 1) * ldb + n;
  }
  cl->validate(a_n, b_n, c_n);
  return true;
}

void dense_gemm(Context* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t rsA,
                int64_t csA, const double* B, int64_t rsB, int64_t csB, double beta, double* C, int64_t rsC,
                int64_t csC) {
  ContractList cl;
  if (!dense_gemm_list(m, n, k, alpha, rsA, csA, rsB, csB, beta, rsC, csC, &cl)) return;
  DeviceContractList d;
  d.upload(ctx, cl);
  launch_contract(ctx, d, A, B, C);
  ctx->sync();
}

BlockGemmPlan::BlockGemmPlan(Sym sym, const QIndex& ar, const QIndex& ac, QN af, const QIndex& br, const QIndex& bc,
                             QN bf)
    : sym_(sym) {
  a_ = make_block_table(sym, {ar, ac}, {Dir::In, Dir::Out}, af);
  b_ = make_block_table(sym, {br, bc}, {Dir::In, Dir::Out}, bf);
  const QN f = fuse(af, bf);
  // task generation over every (a, b) pair (faer.rs:405-433 / tasks.rs:36-81)
  struct Task { int64_t a, b, flops; uint64_t ka, kb; };
  std::map<uint64_t, std::vector<Task>> by_out;
  std::map<uint64_t, std::pair<int, int>> out_secs;
  for (int64_t ia = 0; ia < a_.n_blocks(); ++ia)
    for (int64_t ib = 0; ib < b_.n_blocks(); ++ib) {
      const QN qa1 = ac.q[a_.s(ia, 1)], qb0 = br.q[b_.s(ib, 0)];
      if (qa1 != qb0) continue;
      const QN qr = ar.q[a_.s(ia, 0)], qc = bc.q[b_.s(ib, 1)];
      if (fuse(qr, dual(qc)) != f) continue;
      if (a_.dim(ia, 1) != b_.dim(ib, 0)) throw ArgError("block_gemm: contracted sector dimensions differ");
      QN qs[2] = {qr, qc};
      const uint64_t key = pack_key(sym, qs, 2);
      by_out[key].push_back({ia, ib, a_.dim(ia, 0) * b_.dim(ib, 1) * a_.dim(ia, 1), a_.keys[ia], b_.keys[ib]});
      out_secs[key] = {a_.s(ia, 0), b_.s(ib, 1)};
    }
  // output table: blocks that receive at least one task, ascending key
  int64_t off = 0;
  ContractList& cl = list_;
  cl.flags = 0;
  for (auto& [key, tasks] : by_out) {
    const auto [rs, cs] = out_secs[key];
    const int64_t m = ar.dim[rs], n = bc.dim[cs];
    keys_.push_back(key);
    offsets_.push_back(off);
    // deterministic accumulation order: flops descending, then key_a, key_b
    std::stable_sort(tasks.begin(), tasks.end(), [](const Task& x, const Task& y) {
      if (x.flops != y.flops) return x.flops > y.flops;
      if (x.ka != y.ka) return x.ka < y.ka;
      return x.kb < y.kb;
    });
    cl.begin_out(off, int(m), int(n), int(n));
    for (const Task& t : tasks) {
      const int64_t kk = a_.dim(t.a, 1);
      cl.begin_term(b_.offset[t.b], int(n), int(kk));
      cl.add_src(a_.offset[t.a], int(kk), 1.0);
      flops_ += 2.0 * double(t.flops);
    }
    off += m * n;
  }
  total_ = off;
  cl.finalize();
  cl.validate(a_.total, b_.total, total_);
}

void BlockGemmPlan::run(Context* ctx, const double* a, const double* b, double* c) const {
  if (list_.outs.empty()) return;
  DeviceContractList d;
  d.upload(ctx, list_);
  launch_contract(ctx, d, a, b, c);
  ctx->sync();
}

void block_spmv(Context* ctx, Sym sym, const QIndex& rows, const QIndex& cols, QN flux, const double* a,
                const double* x, double* y) {
  BlockTable t = make_block_table(sym, {rows, cols}, {Dir::In, Dir::Out}, flux);
  ContractList cl;
  cl.flags = 0;
  for (int r = 0; r < rows.n(); ++r) {
    cl.begin_out(rows.off[r], int(rows.dim[r]), 1, 1);
    for (int64_t b = 0; b < t.n_blocks(); ++b) {
      if (t.s(b, 0) != r) continue;
      const int c = t.s(b, 1);
      cl.begin_term(cols.off[c], 1, int(cols.dim[c]));
      cl.add_src(t.offset[b], int(cols.dim[c]), 1.0);
    }
  }
  cl.finalize();
  cl.validate(t.total, cols.total, rows.total);
  DeviceContractList d;
  d.upload(ctx, cl);
  launch_contract(ctx, d, a, x, y);
  ctx->sync();
}

SvdOut svd_truncated(Context* ctx, int64_t m, int64_t n, const double* A, int64_t max_rank, double cutoff,
                     double* U, double* Vt) {
  // the truncate_svd! rule (faer.rs:135-147): keep while count < max_rank and
  // sigma >= cutoff * sigma_max; no minimum, no floor, no degeneracy rule
  TruncRule rule;
  rule.max_bond = std::max<int64_t>(0, max_rank);
  rule.cutoff = cutoff;
  DenseSvd d = dense_svd_truncate(ctx, m, n, A, rule, U, Vt);
  SvdOut out;
  out.rank = d.rank;
  out.s = d.s;
  out.truncation_error = d.sel.abs_error();  // SvdResult: sum of discarded sigma^2
  return out;
}

TruncateOut truncate_svd(Context* ctx, int64_t rows, int64_t cols, const double* A, int64_t max_bond_dim,
                         double svd_cutoff, int64_t min_bond_dim, double* U, double* Vt) {
  TruncRule rule;
  rule.max_bond = std::max<int64_t>(0, max_bond_dim);
  rule.cutoff = svd_cutoff;
  rule.min_bond = min_bond_dim;
  DenseSvd d = dense_svd_truncate(ctx, rows, cols, A, rule, U, Vt);
  TruncateOut out;
  out.bond_dim_new = d.rank;
  out.n_discarded = d.sel.discarded;
  out.s = d.s;
  out.truncation_error = d.sel.rel_error();
  return out;
}

std::vector<double> eigh(Context* ctx, int64_t n, const double* A, double* evecs) {
  if (n <= 0 || n > kSymEigMaxN) throw ArgError("eigh: dimension outside 1..2048");
  DBuf<double> G(ctx, size_t(n * n)), W(ctx, size_t(n));
  TK_HIP(hipMemcpyAsync(G.get(), A, size_t(n * n) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  SymEigBatch eig(ctx);
  eig.values({SymJob{0, 0, int32_t(n)}}, G.get(), n * n, W.get(), n, false);
  eig.vectors({n}, evecs, n * n);
  std::vector<double> w(static_cast<size_t>(n));
  W.download(w.data(), w.size());
  ctx->sync();
  return w;
}

std::vector<double> eigh_lowest(Context* ctx, int64_t n, const double* A, int64_t k, double* evecs) {
  if (n <= 0 || n > kSymEigMaxN) throw ArgError("eigh: dimension outside 1..2048");
  if (k < 0 || k > n) throw ArgError("eigh: k = " + std::to_string(k) + " exceeds the dimension n = " +
                                     std::to_string(n) + " (EighKTooLarge)");
  if (k == 0) return {};
  DBuf<double> G(ctx, size_t(n * n)), W(ctx, size_t(n)), Z(ctx, size_t(n * n));
  TK_HIP(hipMemcpyAsync(G.get(), A, size_t(n * n) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  SymEigBatch eig(ctx);
  eig.values({SymJob{0, 0, int32_t(n)}}, G.get(), n * n, W.get(), n, false);
  eig.vectors({k}, Z.get(), n * n);
  launch_copies(ctx, {CopyJob{0, 0, int32_t(n), int32_t(k), int32_t(n), int32_t(k), 0, 0}}, Z.get(), n * n, evecs,
                n * k);
  std::vector<double> w(static_cast<size_t>(n));
  W.download(w.data(), w.size());
  ctx->sync();
  w.resize(size_t(k));
  return w;
}

namespace {
// Ud (m x k) = U diag(f)
__global__ void scale_columns_kernel(const double* __restrict__ U, const double* __restrict__ f, int64_t m,
                                     int64_t k, double* __restrict__ Ud) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m * k; e += int64_t(gridDim.x) * blockDim.x)
    Ud[e] = U[e] * f[e % k];
}
}  // namespace

void regularized_svd_inverse(Context* ctx, int64_t m, int64_t n, int64_t k, const double* s, const double* U,
                             const double* Vt, double delta, double* out) {
  if (m <= 0 || n <= 0 || k <= 0) throw ArgError("regularized inverse: empty factors");
  if (!(delta > 0)) throw ArgError("regularized inverse: delta must be positive");
  const double d2 = delta * delta;
  std::vector<double> f(static_cast<size_t>(k));
  for (int64_t i = 0; i < k; ++i) f[size_t(i)] = s[i] / (s[i] * s[i] + d2);
  DBuf<double> fd(ctx, size_t(k)), Ud(ctx, size_t(m * k));
  fd.upload(f.data(), f.size());
  const int grid = int(std::max<int64_t>(1, std::min<int64_t>((m * k + 255) / 256, int64_t(ctx->n_cu) * 8)));
  hipLaunchKernelGGL(scale_columns_kernel, dim3(grid), dim3(256), 0, ctx->stream, U, fd.get(), m, k, Ud.get());
  TK_HIP(hipGetLastError());
  // out (n x m) = Vt^T (U diag(f))^T
  dense_gemm(ctx, n, m, k, 1.0, Vt, 1, n, Ud.get(), 1, k, 0.0, out, m, 1);
  ctx->sync();
}

}  // namespace tk
