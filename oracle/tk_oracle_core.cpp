// FP64 CPU oracle: sector tables, block_gemm, H_eff (test infrastructure only;
// see tk_oracle.hpp for what each function restates and how it is pinned).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <vector>

#include "tk_oracle.hpp"
#include "tk_oracle_internal.hpp"

namespace orc {

uint64_t pack1(int sym, Q q) {
  auto b = [](int v) -> uint64_t {
    if (v < -128 || v > 127) throw std::runtime_error("qn out of 8-bit range");
    return uint64_t(v & 0xFF);
  };
  if (sym == 1) return b(q.a);
  return b(q.a) | (b(q.b) << 8);
}

uint64_t packk(int sym, const std::vector<Q>& qs) {
  int w = sym == 1 ? 8 : 16;
  if (int(qs.size()) * w > 64) throw std::runtime_error("key overflow");
  uint64_t k = 0;
  for (size_t i = 0; i < qs.size(); ++i) k |= pack1(sym, qs[i]) << (i * w);
  return k;
}

Leg make_leg(const or_qindex& q) {
  Leg l;
  std::vector<std::pair<Q, long>> v;
  for (int i = 0; i < q.n_sectors; ++i) v.push_back({Q{q.qn[2 * i], q.qn[2 * i + 1]}, long(q.dims[i])});
  // caller order -> sorted order, keeping the caller's full-index map
  std::vector<int> order(v.size());
  for (size_t i = 0; i < v.size(); ++i) order[i] = int(i);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return v[x].first < v[y].first; });
  long off = 0;
  for (int i : order) {
    l.q.push_back(v[i].first);
    l.dim.push_back(v[i].second);
    l.off.push_back(off);
    off += v[i].second;
  }
  l.total = off;
  // caller full index -> sorted full index
  for (size_t i = 0; i < v.size(); ++i) {
    int pos = int(std::find(order.begin(), order.end(), int(i)) - order.begin());
    for (long x = 0; x < v[i].second; ++x) l.perm.push_back(l.off[pos] + x);
  }
  return l;
}

int Leg::find(Q x) const {
  for (size_t i = 0; i < q.size(); ++i)
    if (q[i] == x) return int(i);
  return -1;
}

Q flux_of(const std::vector<Q>& qs, const std::vector<int>& dirs) {
  Q acc{0, 0};
  for (size_t i = 0; i < qs.size(); ++i) acc = dirs[i] == 0 ? add(acc, qs[i]) : add(acc, neg(qs[i]));
  return acc;
}

Tensor make_tensor(int sym, const std::vector<Leg>& legs, const std::vector<int>& dirs, Q flux) {
  Tensor t;
  t.sym = sym;
  t.legs = legs;
  t.dirs = dirs;
  t.flux = flux;
  const int rank = int(legs.size());
  // brute force: every tuple of sectors in mixed radix, keep the flux-allowed
  std::vector<int> idx(rank, 0);
  while (true) {
    std::vector<Q> qs(rank);
    for (int l = 0; l < rank; ++l) qs[l] = legs[l].q[idx[l]];
    if (flux_of(qs, dirs) == flux) {
      Block b;
      b.sec = idx;
      long n = 1;
      for (int l = 0; l < rank; ++l) {
        b.dims.push_back(legs[l].dim[idx[l]]);
        n *= legs[l].dim[idx[l]];
      }
      b.data.assign(size_t(n), 0.0);
      t.blocks[packk(sym, qs)] = b;
    }
    int l = rank - 1;
    while (l >= 0 && ++idx[l] == legs[l].n()) idx[l--] = 0;
    if (l < 0) break;
  }
  return t;
}

long Tensor::total() const {
  long n = 0;
  for (auto& [k, b] : blocks) n += long(b.data.size());
  return n;
}
void Tensor::load(const double* flat) {
  long off = 0;
  for (auto& [k, b] : blocks) {
    std::memcpy(b.data.data(), flat + off, b.data.size() * sizeof(double));
    off += long(b.data.size());
  }
}
void Tensor::store(double* flat) const {
  long off = 0;
  for (auto& [k, b] : blocks) {
    std::memcpy(flat + off, b.data.data(), b.data.size() * sizeof(double));
    off += long(b.data.size());
  }
}
const Block* Tensor::get(const std::vector<int>& sec) const {
  std::vector<Q> qs;
  for (size_t l = 0; l < sec.size(); ++l) qs.push_back(legs[l].q[sec[l]]);
  auto it = blocks.find(packk(sym, qs));
  return it == blocks.end() ? nullptr : &it->second;
}
Block* Tensor::get(const std::vector<int>& sec) {
  return const_cast<Block*>(static_cast<const Tensor*>(this)->get(sec));
}

// Dense MPO site in sorted full-index order.
Mpo make_mpo(const Leg& phys, const Leg& wl, const Leg& wr, const double* w) {
  Mpo m;
  m.phys = phys;
  m.wl = wl;
  m.wr = wr;
  long d = phys.total, a = wl.total, b = wr.total;
  m.w.assign(size_t(d * d * a * b), 0.0);
  for (long si = 0; si < d; ++si)
    for (long so = 0; so < d; ++so)
      for (long x = 0; x < a; ++x)
        for (long y = 0; y < b; ++y)
          m.w[size_t(((phys.perm[si] * d + phys.perm[so]) * a + wl.perm[x]) * b + wr.perm[y])] =
              w[((si * d + so) * a + x) * b + y];
  return m;
}

// Sector position and in-sector offset of a full (sorted) index.
void locate(const Leg& l, long full, int* sec, long* within) {
  for (int s = 0; s < l.n(); ++s)
    if (full < l.off[s] + l.dim[s]) {
      *sec = s;
      *within = full - l.off[s];
      return;
    }
  throw std::runtime_error("index out of range");
}

Tensor make_env(int sym, const Leg& bond, const Leg& w) {
  return make_tensor(sym, {bond, w, bond}, {0, 0, 1}, Q{0, 0});
}

Tensor make_psi2(int sym, const Leg& pi, const Leg& pj, const Leg& bl, const Leg& br) {
  return make_tensor(sym, {pi, pj, bl, br}, {1, 1, 1, 0}, Q{0, 0});
}

// C(m x n) += alpha * A(m x k, lda) * B(k x n, ldb)      (plain loops)
void gemm_nn(long m, long n, long k, double alpha, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc) {
  for (long i = 0; i < m; ++i)
    for (long p = 0; p < k; ++p) {
      const double a = alpha * A[i * lda + p];
      if (a == 0.0) continue;
      const double* b = B + p * ldb;
      double* c = C + i * ldc;
      for (long j = 0; j < n; ++j) c[j] += a * b[j];
    }
}
// C(m x n) += alpha * A(m x k) * B(n x k)^T
void gemm_nt(long m, long n, long k, double alpha, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc) {
  for (long i = 0; i < m; ++i)
    for (long j = 0; j < n; ++j) {
      double s = 0.0;
      const double* a = A + i * lda;
      const double* b = B + j * ldb;
      for (long p = 0; p < k; ++p) s += a[p] * b[p];
      C[i * ldc + j] += alpha * s;
    }
}
// C(m x n) += alpha * A(k x m)^T * B(k x n)
void gemm_tn(long m, long n, long k, double alpha, const double* A, long lda, const double* B, long ldb, double* C,
             long ldc) {
  for (long p = 0; p < k; ++p)
    for (long i = 0; i < m; ++i) {
      const double a = alpha * A[p * lda + i];
      if (a == 0.0) continue;
      const double* b = B + p * ldb;
      double* c = C + i * ldc;
      for (long j = 0; j < n; ++j) c[j] += a * b[j];
    }
}

// y = H_eff x on block-sparse tensors.  For every output block, every W_i
// entry with s_out = s'_i and every W_j entry with s_out = s'_j sharing w_m:
//   y[s'_i s'_j](a'_L, a'_R) += W_i W_j * L_{w_l}(a'_L, a_L) psi[s_i s_j](a_L, a_R) R_{w_r}(a'_R, a_R)
void heff_apply(const Tensor& Lt, const Tensor& Rt, const Mpo& wi, const Mpo& wj, const Tensor& x, Tensor& y,
                int nthreads) {
  const Leg& bl = x.legs[2];
  const Leg& br = x.legs[3];
  const long di = wi.phys.total, dj = wj.phys.total;
  const long nwl = wi.wl.total, nwm = wi.wr.total, nwr = wj.wr.total;
  std::vector<Block*> outs;
  for (auto& [k, b] : y.blocks) outs.push_back(&b);
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (long ob = 0; ob < long(outs.size()); ++ob) {
    Block& yb = *outs[ob];
    std::fill(yb.data.begin(), yb.data.end(), 0.0);
    const int s_oi = yb.sec[0], s_oj = yb.sec[1], s_lp = yb.sec[2], s_rp = yb.sec[3];
    const long dlp = bl.dim[s_lp], drp = br.dim[s_rp];
    std::vector<double> tmp, tmp2;
    for (long xoi = 0; xoi < wi.phys.dim[s_oi]; ++xoi)
      for (long xoj = 0; xoj < wj.phys.dim[s_oj]; ++xoj) {
        const long poi = wi.phys.off[s_oi] + xoi, poj = wj.phys.off[s_oj] + xoj;
        double* yslice = yb.data.data() + (xoi * wj.phys.dim[s_oj] + xoj) * dlp * drp;
        for (long si = 0; si < di; ++si)
          for (long wl = 0; wl < nwl; ++wl)
            for (long wm = 0; wm < nwm; ++wm) {
              const double a = wi.at(si, poi, wl, wm);
              if (a == 0.0) continue;
              for (long sj = 0; sj < dj; ++sj)
                for (long wr = 0; wr < nwr; ++wr) {
                  const double b = wj.at(sj, poj, wm, wr);
                  if (b == 0.0) continue;
                  int ssi, ssj, swl, swr;
                  long xsi, xsj, xwl, xwr;
                  locate(wi.phys, si, &ssi, &xsi);
                  locate(wj.phys, sj, &ssj, &xsj);
                  locate(wi.wl, wl, &swl, &xwl);
                  locate(wj.wr, wr, &swr, &xwr);
                  const Q ql = add(bl.q[s_lp], wi.wl.q[swl]);
                  const Q qr = add(br.q[s_rp], wj.wr.q[swr]);
                  const int sl = bl.find(ql), sr = br.find(qr);
                  if (sl < 0 || sr < 0) continue;
                  const Block* pb = x.get({ssi, ssj, sl, sr});
                  const Block* lb = Lt.get({s_lp, swl, sl});
                  const Block* rb = Rt.get({s_rp, swr, sr});
                  if (!pb || !lb || !rb) continue;
                  const long dl = bl.dim[sl], dr = br.dim[sr];
                  const long dwl = wi.wl.dim[swl], dwr = wj.wr.dim[swr];
                  const double* ps = pb->data.data() + (xsi * pb->dims[1] + xsj) * dl * dr;
                  const double* ls = lb->data.data() + xwl * dl;     // (a'_L, a_L), ld dwl*dl
                  const double* rs = rb->data.data() + xwr * dr;     // (a'_R, a_R), ld dwr*dr
                  tmp.assign(size_t(dlp * dr), 0.0);
                  gemm_nn(dlp, dr, dl, 1.0, ls, dwl * dl, ps, dr, tmp.data(), dr);
                  gemm_nt(dlp, drp, dr, a * b, tmp.data(), dr, rs, dwr * dr, yslice, drp);
                }
            }
      }
  }
}

void heff_diag(const Tensor& Lt, const Tensor& Rt, const Mpo& wi, const Mpo& wj, Tensor& d) {
  const Leg& bl = d.legs[2];
  const Leg& br = d.legs[3];
  const long nwl = wi.wl.total, nwm = wi.wr.total, nwr = wj.wr.total;
  for (auto& [k, db] : d.blocks) {
    std::fill(db.data.begin(), db.data.end(), 0.0);
    const int s_i = db.sec[0], s_j = db.sec[1], s_l = db.sec[2], s_r = db.sec[3];
    const long dl = bl.dim[s_l], dr = br.dim[s_r];
    for (long xi = 0; xi < wi.phys.dim[s_i]; ++xi)
      for (long xj = 0; xj < wj.phys.dim[s_j]; ++xj) {
        const long pi = wi.phys.off[s_i] + xi, pj = wj.phys.off[s_j] + xj;
        double* ds = db.data.data() + (xi * wj.phys.dim[s_j] + xj) * dl * dr;
        for (long wl = 0; wl < nwl; ++wl)
          for (long wm = 0; wm < nwm; ++wm) {
            const double a = wi.at(pi, pi, wl, wm);
            if (a == 0.0) continue;
            for (long wr = 0; wr < nwr; ++wr) {
              const double b = wj.at(pj, pj, wm, wr);
              if (b == 0.0) continue;
              int swl, swr;
              long xwl, xwr;
              locate(wi.wl, wl, &swl, &xwl);
              locate(wj.wr, wr, &swr, &xwr);
              const Block* lb = Lt.get({s_l, swl, s_l});
              const Block* rb = Rt.get({s_r, swr, s_r});
              if (!lb || !rb) continue;
              const long dwl = wi.wl.dim[swl], dwr = wj.wr.dim[swr];
              for (long p = 0; p < dl; ++p)
                for (long q = 0; q < dr; ++q)
                  ds[p * dr + q] += a * b * lb->data[p * dwl * dl + xwl * dl + p] * rb->data[q * dwr * dr + xwr * dr + q];
            }
          }
      }
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

int or_pack_key(int32_t sym, int32_t rank, const int32_t* qn, uint64_t* key) {
  try {
    std::vector<Q> qs;
    for (int i = 0; i < rank; ++i) qs.push_back(Q{qn[2 * i], qn[2 * i + 1]});
    *key = packk(sym, qs);
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_block_table(int32_t sym, int32_t rank, const or_qindex* legs, const int32_t* dirs, const int32_t* flux,
                   int64_t capacity, int64_t* n_blocks, uint64_t* keys, int64_t* offsets, int64_t* total) {
  try {
    std::vector<Leg> ls;
    std::vector<int> ds;
    for (int i = 0; i < rank; ++i) {
      ls.push_back(make_leg(legs[i]));
      ds.push_back(dirs[i]);
    }
    Tensor t = make_tensor(sym, ls, ds, Q{flux[0], flux[1]});
    *n_blocks = int64_t(t.blocks.size());
    if (total) *total = t.total();
    if (int64_t(t.blocks.size()) > capacity) return 7;
    long off = 0, i = 0;
    for (auto& [k, b] : t.blocks) {
      if (keys) keys[i] = k;
      if (offsets) offsets[i] = off;
      off += long(b.data.size());
      ++i;
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_fuse_map(int32_t sym, int32_t n_legs, const or_qindex* legs, const int32_t* dirs, int64_t capacity,
                int32_t* n_fused, int32_t* fused_qn, int64_t* fused_dims, int64_t* n_combos,
                int32_t* combo_fused_sector, int64_t* combo_offset) {
  (void)sym;
  try {
    std::vector<Leg> ls;
    for (int i = 0; i < n_legs; ++i) ls.push_back(make_leg(legs[i]));
    // combos with later leg fastest: mixed-radix counting
    long ncomb = 1;
    for (auto& l : ls) ncomb *= l.n();
    std::vector<Q> cq(ncomb);
    std::vector<long> cd(ncomb);
    for (long c = 0; c < ncomb; ++c) {
      long r = c;
      Q q{0, 0};
      long d = 1;
      for (int l = n_legs - 1; l >= 0; --l) {
        int s = int(r % ls[l].n());
        r /= ls[l].n();
        q = add(q, dirs[l] == 0 ? ls[l].q[s] : neg(ls[l].q[s]));
        d *= ls[l].dim[s];
      }
      cq[c] = q;
      cd[c] = d;
    }
    std::map<Q, long> tot;
    std::vector<long> off(ncomb);
    for (long c = 0; c < ncomb; ++c) {
      off[c] = tot[cq[c]];
      tot[cq[c]] += cd[c];
    }
    *n_fused = int32_t(tot.size());
    *n_combos = ncomb;
    if (long(tot.size()) > capacity || ncomb > capacity) return 7;
    int i = 0;
    std::map<Q, int> pos;
    for (auto& [q, d] : tot) {
      if (fused_qn) {
        fused_qn[2 * i] = q.a;
        fused_qn[2 * i + 1] = q.b;
      }
      if (fused_dims) fused_dims[i] = d;
      pos[q] = i++;
    }
    for (long c = 0; c < ncomb; ++c) {
      if (combo_fused_sector) combo_fused_sector[c] = pos[cq[c]];
      if (combo_offset) combo_offset[c] = off[c];
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_block_gemm(int32_t sym, const or_qindex* a_rows, const or_qindex* a_cols, const int32_t* a_flux,
                  const double* a_data, const or_qindex* b_rows, const or_qindex* b_cols, const int32_t* b_flux,
                  const double* b_data, int64_t capacity, int64_t* n_blocks, uint64_t* keys, int64_t* offsets,
                  double* c_data, int64_t* c_total) {
  try {
    Leg ar = make_leg(*a_rows), ac = make_leg(*a_cols), br = make_leg(*b_rows), bc = make_leg(*b_cols);
    Tensor A = make_tensor(sym, {ar, ac}, {0, 1}, Q{a_flux[0], a_flux[1]});
    Tensor B = make_tensor(sym, {br, bc}, {0, 1}, Q{b_flux[0], b_flux[1]});
    A.load(a_data);
    B.load(b_data);
    Q f = add(A.flux, B.flux);
    // every (a, b) pair with q_a[1] == q_b[0] and q_row - q_col == flux (tasks.rs:36-81)
    std::map<uint64_t, Block> out;
    for (auto& [ka, ba] : A.blocks)
      for (auto& [kb, bb] : B.blocks) {
        Q qa1 = ac.q[ba.sec[1]], qb0 = br.q[bb.sec[0]];
        if (!(qa1 == qb0)) continue;
        Q qr = ar.q[ba.sec[0]], qc = bc.q[bb.sec[1]];
        if (!(add(qr, neg(qc)) == f)) continue;
        uint64_t k = packk(sym, {qr, qc});
        Block& ob = out[k];
        if (ob.data.empty()) {
          ob.sec = {ba.sec[0], bb.sec[1]};
          ob.dims = {ba.dims[0], bb.dims[1]};
          ob.data.assign(size_t(ba.dims[0] * bb.dims[1]), 0.0);
        }
        gemm_nn(ba.dims[0], bb.dims[1], ba.dims[1], 1.0, ba.data.data(), ba.dims[1], bb.data.data(), bb.dims[1],
                ob.data.data(), bb.dims[1]);
      }
    *n_blocks = int64_t(out.size());
    long off = 0, i = 0;
    for (auto& [k, b] : out) {
      if (i < capacity) {
        if (keys) keys[i] = k;
        if (offsets) offsets[i] = off;
        if (c_data) std::memcpy(c_data + off, b.data.data(), b.data.size() * sizeof(double));
      }
      off += long(b.data.size());
      ++i;
    }
    if (c_total) *c_total = off;
    return int64_t(out.size()) > capacity ? 7 : 0;
  } catch (...) {
    return 1;
  }
}

int or_heff_apply(int32_t sym, const or_qindex* phys_i, const or_qindex* phys_j, const or_qindex* bond_l,
                  const or_qindex* bond_r, const or_qindex* w_l, const or_qindex* w_m, const or_qindex* w_r,
                  const double* wi, const double* wj, const double* L, const double* R, const double* x, double* y,
                  int32_t nthreads) {
  try {
    Leg pi = make_leg(*phys_i), pj = make_leg(*phys_j), bl = make_leg(*bond_l), br = make_leg(*bond_r);
    Leg wlg = make_leg(*w_l), wmg = make_leg(*w_m), wrg = make_leg(*w_r);
    Mpo mi = make_mpo(pi, wlg, wmg, wi), mj = make_mpo(pj, wmg, wrg, wj);
    Tensor Lt = make_env(sym, bl, wlg), Rt = make_env(sym, br, wrg);
    Lt.load(L);
    Rt.load(R);
    Tensor xt = make_psi2(sym, pi, pj, bl, br);
    xt.load(x);
    Tensor yt = xt;
    heff_apply(Lt, Rt, mi, mj, xt, yt, nthreads);
    yt.store(y);
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_heff_diag(int32_t sym, const or_qindex* phys_i, const or_qindex* phys_j, const or_qindex* bond_l,
                 const or_qindex* bond_r, const or_qindex* w_l, const or_qindex* w_m, const or_qindex* w_r,
                 const double* wi, const double* wj, const double* L, const double* R, double* d) {
  try {
    Leg pi = make_leg(*phys_i), pj = make_leg(*phys_j), bl = make_leg(*bond_l), br = make_leg(*bond_r);
    Leg wlg = make_leg(*w_l), wmg = make_leg(*w_m), wrg = make_leg(*w_r);
    Mpo mi = make_mpo(pi, wlg, wmg, wi), mj = make_mpo(pj, wmg, wrg, wj);
    Tensor Lt = make_env(sym, bl, wlg), Rt = make_env(sym, br, wrg);
    Lt.load(L);
    Rt.load(R);
    Tensor dt = make_psi2(sym, pi, pj, bl, br);
    heff_diag(Lt, Rt, mi, mj, dt);
    dt.store(d);
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_heff_dense_reference(int32_t sym, const or_qindex* phys_i, const or_qindex* phys_j, const or_qindex* bond_l,
                            const or_qindex* bond_r, const or_qindex* w_l, const or_qindex* w_m,
                            const or_qindex* w_r, const double* wi, const double* wj, const double* L,
                            const double* R, const double* x_dense, double* y_dense) {
  try {
    Leg pi = make_leg(*phys_i), pj = make_leg(*phys_j), bl = make_leg(*bond_l), br = make_leg(*bond_r);
    Leg wlg = make_leg(*w_l), wmg = make_leg(*w_m), wrg = make_leg(*w_r);
    Mpo mi = make_mpo(pi, wlg, wmg, wi), mj = make_mpo(pj, wmg, wrg, wj);
    Tensor Lt = make_env(sym, bl, wlg), Rt = make_env(sym, br, wrg);
    Lt.load(L);
    Rt.load(R);
    // Densify the environments as the reference stores them: (D_bra, w, D_ket),
    // flat a_bra*dw*dk + w*dk + a_ket (environments/mod.rs:90-95).
    auto densify = [](const Tensor& t, const Leg& b, const Leg& w) {
      std::vector<double> e(size_t(b.total * w.total * b.total), 0.0);
      for (auto& [k, blk] : t.blocks) {
        long ob = b.off[blk.sec[0]], ow = w.off[blk.sec[1]], ok = b.off[blk.sec[2]];
        for (long p = 0; p < blk.dims[0]; ++p)
          for (long q = 0; q < blk.dims[1]; ++q)
            for (long r = 0; r < blk.dims[2]; ++r)
              e[size_t(((ob + p) * w.total + ow + q) * b.total + ok + r)] =
                  blk.data[size_t((p * blk.dims[1] + q) * blk.dims[2] + r)];
      }
      return e;
    };
    std::vector<double> el = densify(Lt, bl, wlg), er = densify(Rt, br, wrg);
    const long d_i = pi.total, d_j = pj.total, d_l = bl.total, d_r = br.total;
    const long dim = d_i * d_j * d_l * d_r;
    if (dim > 6000) return 7;
    auto env_get = [](const std::vector<double>& e, long dw, long dk, long a, long w, long k) {
      return e[size_t((a * dw + w) * dk + k)];
    };
    // MPO blocks: every flux-allowed (s_in, s_out, w_l, w_r) sector tuple of
    // the dense W, iterated as BlockSparseTensor::iter_blocks would (key order).
    Tensor Wi = make_tensor(sym, {pi, pi, wlg, wmg}, {0, 1, 0, 1}, Q{0, 0});
    Tensor Wj = make_tensor(sym, {pj, pj, wmg, wrg}, {0, 1, 0, 1}, Q{0, 0});
    auto fill = [](Tensor& T, const Mpo& m) {
      for (auto& [k, b] : T.blocks) {
        long o0 = T.legs[0].off[b.sec[0]], o1 = T.legs[1].off[b.sec[1]], o2 = T.legs[2].off[b.sec[2]],
             o3 = T.legs[3].off[b.sec[3]];
        for (long a = 0; a < b.dims[0]; ++a)
          for (long c = 0; c < b.dims[1]; ++c)
            for (long e = 0; e < b.dims[2]; ++e)
              for (long f = 0; f < b.dims[3]; ++f)
                b.data[size_t(((a * b.dims[1] + c) * b.dims[2] + e) * b.dims[3] + f)] = m.at(o0 + a, o1 + c, o2 + e, o3 + f);
      }
    };
    fill(Wi, mi);
    fill(Wj, mj);
    std::vector<double> heff(size_t(dim * dim), 0.0);
    const long wl_tot = wlg.total, wr_tot = wrg.total;
    // environments/mod.rs:510-590, loop for loop
    for (auto& [ki, bi] : Wi.blocks) {
      const long d_sin_i = bi.dims[0], d_sout_i = bi.dims[1], d_wl_i = bi.dims[2], d_wm_i = bi.dims[3];
      const long sin_i_off = pi.off[bi.sec[0]], sout_i_off = pi.off[bi.sec[1]], wl_off = wlg.off[bi.sec[2]];
      for (auto& [kj, bj] : Wj.blocks) {
        const long d_sin_j = bj.dims[0], d_sout_j = bj.dims[1], d_wm_j = bj.dims[2], d_wr_j = bj.dims[3];
        if (!(wmg.q[bi.sec[3]] == wmg.q[bj.sec[2]])) continue;
        const long sin_j_off = pj.off[bj.sec[0]], sout_j_off = pj.off[bj.sec[1]], wr_off = wrg.off[bj.sec[3]];
        for (long s_in_i = 0; s_in_i < d_sin_i; ++s_in_i)
          for (long s_out_i = 0; s_out_i < d_sout_i; ++s_out_i)
            for (long s_in_j = 0; s_in_j < d_sin_j; ++s_in_j)
              for (long s_out_j = 0; s_out_j < d_sout_j; ++s_out_j)
                for (long wl = 0; wl < d_wl_i; ++wl)
                  for (long wm = 0; wm < std::min(d_wm_i, d_wm_j); ++wm) {
                    const double wi_val = bi.data[size_t(s_in_i * d_sout_i * d_wl_i * d_wm_i + s_out_i * d_wl_i * d_wm_i + wl * d_wm_i + wm)];
                    for (long wr = 0; wr < d_wr_j; ++wr) {
                      const double wj_val = bj.data[size_t(s_in_j * d_sout_j * d_wm_j * d_wr_j + s_out_j * d_wm_j * d_wr_j + wm * d_wr_j + wr)];
                      const double w_contrib = wi_val * wj_val;
                      for (long a_l = 0; a_l < d_l; ++a_l)
                        for (long ap_l = 0; ap_l < d_l; ++ap_l) {
                          const double lw = env_get(el, wl_tot, d_l, ap_l, wl_off + wl, a_l) * w_contrib;
                          for (long a_r = 0; a_r < d_r; ++a_r)
                            for (long ap_r = 0; ap_r < d_r; ++ap_r) {
                              const double r_val = env_get(er, wr_tot, d_r, ap_r, wr_off + wr, a_r);
                              const long in_idx = (sin_i_off + s_in_i) * d_j * d_l * d_r +
                                                  (sin_j_off + s_in_j) * d_l * d_r + a_l * d_r + a_r;
                              const long out_idx = (sout_i_off + s_out_i) * d_j * d_l * d_r +
                                                   (sout_j_off + s_out_j) * d_l * d_r + ap_l * d_r + ap_r;
                              heff[size_t(out_idx * dim + in_idx)] += lw * r_val;
                            }
                        }
                    }
                  }
      }
    }
    for (long i = 0; i < dim; ++i) {
      double s = 0.0;
      for (long j = 0; j < dim; ++j) s += heff[size_t(i * dim + j)] * x_dense[j];
      y_dense[i] = s;
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"

extern "C" int or_qindex_layout(const or_qindex* q, int32_t* qn_out, int64_t* dims_out, int64_t* offsets_out,
                                int64_t* total_out) {
  try {
    orc::Leg l = orc::make_leg(*q);
    for (int i = 0; i < l.n(); ++i) {
      qn_out[2 * i] = l.q[i].a;
      qn_out[2 * i + 1] = l.q[i].b;
      dims_out[i] = l.dim[i];
      offsets_out[i] = l.off[i];
    }
    *total_out = l.total;
    return 0;
  } catch (...) {
    return 1;
  }
}
