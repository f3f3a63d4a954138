// Internal types of the CPU oracle (test infrastructure only).
#pragma once
#include <cstdint>
#include <map>
#include <vector>

#include "tk_oracle.hpp"

namespace orc {

struct Q {
  int a = 0, b = 0;
  bool operator==(const Q& o) const { return a == o.a && b == o.b; }
  bool operator<(const Q& o) const { return a < o.a || (a == o.a && b < o.b); }
};
inline Q add(Q x, Q y) { return Q{x.a + y.a, x.b + y.b}; }
inline Q neg(Q x) { return Q{-x.a, -x.b}; }

struct Leg {
  std::vector<Q> q;
  std::vector<long> dim, off;
  std::vector<long> perm;  // caller full index -> sorted full index
  long total = 0;
  int n() const { return int(q.size()); }
  int find(Q x) const;
};
Leg make_leg(const or_qindex& q);
// Leg from (q, dim) pairs, sorted.
inline Leg make_leg_v(const std::vector<std::pair<Q, long>>& v) {
  std::vector<int32_t> qn;
  std::vector<int64_t> d;
  for (auto& [q, n] : v) {
    qn.push_back(q.a);
    qn.push_back(q.b);
    d.push_back(n);
  }
  or_qindex oq{int32_t(v.size()), qn.data(), d.data()};
  return make_leg(oq);
}

struct Block {
  std::vector<int> sec;
  std::vector<long> dims;
  std::vector<double> data;
};

struct Tensor {
  int sym = 1;
  std::vector<Leg> legs;
  std::vector<int> dirs;  // 0 = In, 1 = Out
  Q flux;
  std::map<uint64_t, Block> blocks;  // ascending key
  long total() const;
  void load(const double* flat);
  void store(double* flat) const;
  const Block* get(const std::vector<int>& sec) const;
  Block* get(const std::vector<int>& sec);
};

struct Mpo {
  Leg phys, wl, wr;
  std::vector<double> w;
  double at(long si, long so, long a, long b) const {
    return w[size_t(((si * phys.total + so) * wl.total + a) * wr.total + b)];
  }
};

uint64_t pack1(int sym, Q q);
uint64_t packk(int sym, const std::vector<Q>& qs);
Q flux_of(const std::vector<Q>& qs, const std::vector<int>& dirs);
Tensor make_tensor(int sym, const std::vector<Leg>& legs, const std::vector<int>& dirs, Q flux);
Mpo make_mpo(const Leg& phys, const Leg& wl, const Leg& wr, const double* w);
void locate(const Leg& l, long full, int* sec, long* within);
Tensor make_env(int sym, const Leg& bond, const Leg& w);
Tensor make_psi2(int sym, const Leg& pi, const Leg& pj, const Leg& bl, const Leg& br);
void gemm_nn(long m, long n, long k, double alpha, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc);
void gemm_nt(long m, long n, long k, double alpha, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc);
void gemm_tn(long m, long n, long k, double alpha, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc);
void heff_apply(const Tensor& Lt, const Tensor& Rt, const Mpo& wi, const Mpo& wj, const Tensor& x, Tensor& y,
                int nthreads);
void heff_diag(const Tensor& Lt, const Tensor& Rt, const Mpo& wi, const Mpo& wj, Tensor& d);

}  // namespace orc
