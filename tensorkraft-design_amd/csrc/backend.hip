// Dense and rank-2 block-sparse backend operations of the LinAlgBackend /
// SparseLinAlgBackend traits (tk-linalg/src/traits.rs:25-205) on the
// grouped-contraction kernel and the batched Jacobi eigensolver.
#include "backend.hpp"

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

void dense_gemm(Context* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t rsA,
                int64_t csA, const double* B, int64_t rsB, int64_t csB, double beta, double* C, int64_t rsC,
                int64_t csC) {
  if (m < 0 || n < 0 || k < 0) throw ArgError("gemm: negative dimension");
  if (m == 0 || n == 0) return;
  int64_t lda, ldb, ldc;
  bool ta, tb, tc;
  if (!as_slice(m, k, rsA, csA, &lda, &ta) || !as_slice(k, n, rsB, csB, &ldb, &tb) ||
      !as_slice(m, n, rsC, csC, &ldc, &tc))
    throw ArgError("gemm: operand without a unit stride (NonContiguousForBlas)");
  ContractList cl;
  cl.flags = (ta ? kTransA : 0) | (tb ? kTransB : 0) | (tc ? kTransC : 0);
  cl.beta = beta;
  cl.begin_out(0, int(m), int(n), int(ldc));
  if (k > 0 && alpha != 0.0) {
    cl.begin_term(0, int(ldb), int(k));
    cl.add_src(0, int(lda), alpha);
  }
  const int64_t a_n = ta ? (k - 1) * lda + m : (m - 1) * lda + k;
  const int64_t b_n = tb ? (n - 1) * ldb + k : (k - 1) * ldb + n;
  const int64_t c_n = tc ? (n - 1) * ldc + m : (m - 1) * ldc + n;
  if (k == 0 || alpha == 0.0) {
    // C = beta C: an output with no terms under beta scales (beta = 0 zeroes)
    cl.finalize();
    DeviceContractList d;
    d.upload(ctx, cl);
    launch_contract(ctx, d, A, B, C);
    ctx->sync();
    return;
  }
  run(ctx, cl, a_n, b_n, c_n, A, B, C);
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
  if (m <= 0 || n <= 0) throw ArgError("svd: empty matrix");
  const bool rows_side = m < n;  // Gram of the smaller side
  const int64_t g = rows_side ? m : n;
  if (g > kSymEigMaxN) throw ArgError("svd: the smaller dimension exceeds the eigensolver limit");
  DBuf<double> G(ctx, size_t(g * g)), E(ctx, size_t(g * g)), W(ctx, size_t(g));
  if (rows_side)  // G = A A^T
    dense_gemm(ctx, m, m, n, 1.0, A, n, 1, A, 1, n, 0.0, G.get(), m, 1);
  else  // G = A^T A
    dense_gemm(ctx, n, n, m, 1.0, A, 1, n, A, n, 1, 0.0, G.get(), n, 1);
  SymEigBatch eig(ctx);
  eig.values({SymJob{0, 0, int32_t(g)}}, G.get(), g * g, W.get(), g, true);
  std::vector<double> lam(static_cast<size_t>(g));
  W.download(lam.data(), lam.size());
  ctx->sync();
  SvdOut out;
  const double lmax = std::max(0.0, lam[0]);
  double total = 0;
  for (double l : lam) total += std::max(0.0, l);
  int64_t r = 0;
  // keep while count < max_rank and sigma >= cutoff * sigma_max (faer.rs:135-147);
  // singular values at the Gram noise floor (sigma^2 <= 1e-15 sigma_max^2) count as zero
  while (r < g && r < max_rank && lam[r] > 1e-15 * lmax && std::sqrt(lam[r]) >= cutoff * std::sqrt(lmax)) ++r;
  double kept = 0;
  for (int64_t j = 0; j < r; ++j) {
    out.s.push_back(std::sqrt(lam[j]));
    kept += lam[j];
  }
  out.rank = r;
  out.truncation_error = std::max(0.0, total - kept);  // SvdResult: sum of discarded sigma^2
  if (r == 0) return out;
  eig.vectors({r}, E.get(), g * g);
  // scale columns of the eigenvectors by 1/sigma to form the other factor
  std::vector<double> inv(static_cast<size_t>(r));
  for (int64_t j = 0; j < r; ++j) inv[j] = 1.0 / out.s[j];
  if (rows_side) {
    // U = E[:, :r] (m x r, row-major, ld g); Vt = diag(1/s) U^T A  (r x n)
    for (int64_t i = 0; i < m; ++i)
      TK_HIP(hipMemcpyAsync(U + i * r, E.get() + i * g, size_t(r) * sizeof(double), hipMemcpyDeviceToDevice,
                            ctx->stream));
    dense_gemm(ctx, r, n, m, 1.0, U, 1, r, A, n, 1, 0.0, Vt, n, 1);
    // row j of Vt scaled by 1/s_j: Vt = diag(inv) Vt via a diagonal-matrix product
    DBuf<double> D(ctx, size_t(r * r)), T(ctx, size_t(r * n));
    std::vector<double> dh(size_t(r * r), 0.0);
    for (int64_t j = 0; j < r; ++j) dh[size_t(j * r + j)] = inv[j];
    D.upload(dh.data(), dh.size());
    TK_HIP(hipMemcpyAsync(T.get(), Vt, size_t(r * n) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    dense_gemm(ctx, r, n, r, 1.0, D.get(), r, 1, T.get(), n, 1, 0.0, Vt, n, 1);
  } else {
    // Vt = E[:, :r]^T (r x n); U = A E[:, :r] diag(1/s) (m x r)
    DBuf<double> V(ctx, size_t(n * r));
    for (int64_t i = 0; i < n; ++i)
      TK_HIP(hipMemcpyAsync(V.get() + i * r, E.get() + i * g, size_t(r) * sizeof(double), hipMemcpyDeviceToDevice,
                            ctx->stream));
    std::vector<double> dh(size_t(r * r), 0.0);
    for (int64_t j = 0; j < r; ++j) dh[size_t(j * r + j)] = inv[j];
    DBuf<double> D(ctx, size_t(r * r)), T(ctx, size_t(m * r));
    D.upload(dh.data(), dh.size());
    dense_gemm(ctx, m, r, n, 1.0, A, n, 1, V.get(), r, 1, 0.0, T.get(), r, 1);
    dense_gemm(ctx, m, r, r, 1.0, T.get(), r, 1, D.get(), r, 1, 0.0, U, r, 1);
    // Vt = I V^T
    std::vector<double> ih(size_t(r * r), 0.0);
    for (int64_t j = 0; j < r; ++j) ih[size_t(j * r + j)] = 1.0;
    DBuf<double> I(ctx, size_t(r * r));
    I.upload(ih.data(), ih.size());
    dense_gemm(ctx, r, n, r, 1.0, I.get(), r, 1, V.get(), 1, r, 0.0, Vt, n, 1);
  }
  ctx->sync();
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

}  // namespace tk
