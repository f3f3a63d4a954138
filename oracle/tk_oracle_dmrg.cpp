// FP64 CPU oracle for two-site DMRG (test infrastructure only; see tk_oracle.hpp).
//
// Restates tk-dmrg's step (sweep/mod.rs:348-448) with the design's defect
// fixes: warm-started Davidson (eigensolver/mod.rs:325-516 with the diagonal
// preconditioner and the reference's classical Jacobi subspace solver,
// eigensolver/mod.rs:532-601), full SVD of the (s_i a_L) x (s_j a_R)
// matricisation per middle charge by one-sided Jacobi (Hestenes), the
// truncation rule of truncation/mod.rs:56-127 (relative cutoff, max bond,
// min bond) plus a numerical-zero floor (sigma^2 <= 1e-26 sigma_max^2) and
// the rule of never cutting inside a group |sigma_k - sigma_k+1| <= tol *
// sigma_max, write-back into the MPS, and block-sparse
// environment updates (environments/mod.rs:225-483).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "tk_oracle.hpp"
#include "tk_oracle_internal.hpp"

namespace orc {

// sigma^2 <= kLambdaZero sigma_max^2 is a numerical zero (sigma <= 1e-13
// sigma_max, below the reference's default cutoff of 1e-12).
constexpr double kLambdaZero = 1e-26;

namespace {

double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// symmetric_eigen_small (eigensolver/mod.rs:532-601): classical Jacobi, largest
// off-diagonal pivot, threshold 1e-15, at most 100 n^2 rotations.  Eigenvectors
// column-major-by-index v[i*n + j] = component i of vector j.
void jacobi_small(const std::vector<double>& mat, int n, std::vector<double>& evals, std::vector<double>& v) {
  // Same rotations in the same order as the reference: the pivot is the first
  // largest |a_ij| (i < j) in row-major order.  Instead of rescanning the
  // triangle and copying the matrix for every rotation, the first maximum of
  // every row is tracked (rm, rc) and only rows p, q and columns p, q are
  // rewritten, from the values before the rotation as in the reference.
  std::vector<double> a = mat;
  v.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
  std::vector<double> rm(size_t(std::max(n, 1)), 0.0);
  std::vector<int> rc(size_t(std::max(n, 1)), -1);
  auto rescan = [&](int i) {
    rm[i] = 0.0;
    rc[i] = -1;
    for (int j = i + 1; j < n; ++j)
      if (std::fabs(a[i * n + j]) > rm[i]) {
        rm[i] = std::fabs(a[i * n + j]);
        rc[i] = j;
      }
  };
  // a changed entry (i, j), i < j, of a row whose maximum is elsewhere
  auto touch = [&](int i, int j) {
    const double x = std::fabs(a[i * n + j]);
    if (x > rm[i] || (x == rm[i] && x > 0.0 && j < rc[i])) {
      rm[i] = x;
      rc[i] = j;
    }
  };
  for (int i = 0; i < n; ++i) rescan(i);
  std::vector<double> colp(size_t(std::max(n, 1))), colq(size_t(std::max(n, 1)));
  for (long it = 0; it < 100L * n * n; ++it) {
    double mx = 0.0;
    int p = 0, q = 1;
    for (int i = 0; i < n; ++i)
      if (rm[i] > mx) {
        mx = rm[i];
        p = i;
        q = rc[i];
      }
    if (mx < 1e-15 || n < 2) break;
    const double app = a[p * n + p], aqq = a[q * n + q], apq = a[p * n + q];
    const double theta = std::fabs(app - aqq) < 1e-30 ? M_PI / 4 : 0.5 * std::atan(2.0 * apq / (app - aqq));
    const double c = std::cos(theta), s = std::sin(theta);
    for (int i = 0; i < n; ++i)
      if (i != p && i != q) {
        colp[i] = c * a[i * n + p] + s * a[i * n + q];
        colq[i] = -s * a[i * n + p] + c * a[i * n + q];
      }
    for (int i = 0; i < n; ++i)
      if (i != p && i != q) {
        a[i * n + p] = a[p * n + i] = colp[i];
        a[i * n + q] = a[q * n + i] = colq[i];
      }
    a[p * n + p] = c * c * app + 2.0 * s * c * apq + s * s * aqq;
    a[q * n + q] = s * s * app - 2.0 * s * c * apq + c * c * aqq;
    a[p * n + q] = 0.0;
    a[q * n + p] = 0.0;
    rescan(p);
    rescan(q);
    for (int i = 0; i < q; ++i) {
      if (i == p) continue;
      if (rc[i] == p || rc[i] == q) {
        rescan(i);
      } else {
        if (i < p) touch(i, p);
        touch(i, q);
      }
    }
    for (int i = 0; i < n; ++i) {
      const double vip = v[i * n + p], viq = v[i * n + q];
      v[i * n + p] = c * vip + s * viq;
      v[i * n + q] = -s * vip + c * viq;
    }
  }
  evals.resize(n);
  for (int i = 0; i < n; ++i) evals[i] = a[i * n + i];
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// One-sided Jacobi SVD of an m x n row-major matrix: A = U diag(s) V^T,
// singular values descending.  U is m x r, V is n x r, r = min(m, n).
void svd_jacobi(const std::vector<double>& A, long m, long n, std::vector<double>& U, std::vector<double>& s,
                std::vector<double>& V) {
  const bool tr = m < n;  // work on the tall orientation
  const long rows = tr ? n : m, cols = tr ? m : n;
  // columns of W (rows x cols), stored column by column
  std::vector<double> W(size_t(rows * cols));
  for (long i = 0; i < rows; ++i)
    for (long j = 0; j < cols; ++j) W[size_t(j * rows + i)] = tr ? A[size_t(j * n + i)] : A[size_t(i * n + j)];
  std::vector<double> Q(size_t(cols * cols), 0.0);  // right rotations, column-major
  for (long j = 0; j < cols; ++j) Q[size_t(j * cols + j)] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (long p = 0; p < cols - 1; ++p)
      for (long q = p + 1; q < cols; ++q) {
        double* wp = &W[size_t(p * rows)];
        double* wq = &W[size_t(q * rows)];
        double alpha = 0, beta = 0, gamma = 0;
        for (long i = 0; i < rows; ++i) {
          alpha += wp[i] * wp[i];
          beta += wq[i] * wq[i];
          gamma += wp[i] * wq[i];
        }
        if (std::fabs(gamma) <= 1e-15 * std::sqrt(alpha * beta) || gamma == 0.0) continue;
        off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (long i = 0; i < rows; ++i) {
          const double x = wp[i], y = wq[i];
          wp[i] = c * x - sn * y;
          wq[i] = sn * x + c * y;
        }
        double* qp = &Q[size_t(p * cols)];
        double* qq = &Q[size_t(q * cols)];
        for (long i = 0; i < cols; ++i) {
          const double x = qp[i], y = qq[i];
          qp[i] = c * x - sn * y;
          qq[i] = sn * x + c * y;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> sv(cols);
  for (long j = 0; j < cols; ++j) {
    double t = 0;
    for (long i = 0; i < rows; ++i) t += W[size_t(j * rows + i)] * W[size_t(j * rows + i)];
    sv[j] = std::sqrt(t);
  }
  std::vector<long> ord(cols);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return sv[a] > sv[b]; });
  const long r = cols;
  s.resize(r);
  // left vectors of the tall matrix = W columns / s, right vectors = Q columns
  std::vector<double> L(size_t(rows * r)), Rm(size_t(cols * r));
  for (long k = 0; k < r; ++k) {
    const long j = ord[k];
    s[k] = sv[j];
    for (long i = 0; i < rows; ++i) L[size_t(i * r + k)] = sv[j] > 0 ? W[size_t(j * rows + i)] / sv[j] : 0.0;
    for (long i = 0; i < cols; ++i) Rm[size_t(i * r + k)] = Q[size_t(j * cols + i)];
  }
  if (!tr) {
    U.swap(L);
    V.swap(Rm);
  } else {
    U.swap(Rm);
    V.swap(L);
  }
}

// Davidson (eigensolver/mod.rs:325-516) on any symmetric matvec, with the
// caller's start vector and an optional diagonal preconditioner.
double davidson_core(const std::function<void(const std::vector<double>&, std::vector<double>&)>& matvec, long dim,
                     const std::vector<double>* diag, int max_sub_in, int max_iter, double tol,
                     std::vector<double>& x) {
    const int max_sub = int(std::min<long>(max_sub_in, dim));
    std::vector<double> v0 = x;
    double nrm = std::sqrt(dot(v0, v0));
    for (double& e : v0) e /= nrm;
    std::vector<std::vector<double>> basis{v0}, ab(1);
    matvec(basis[0], ab[0]);
    double eig = 0.0;
    std::vector<double> y = basis[0];
    // The reference rebuilds the projected matrix from scratch every iteration
    // (eigensolver/mod.rs:372-384); entries of earlier iterations are the same
    // dot products of unchanged vectors, so they are kept (bit-identical
    // values) and only the new column is formed.  A restart clears them.
    std::vector<std::vector<double>> hc;
    for (int it = 0; it < max_iter; ++it) {
      const int ns = int(basis.size());
      while (int(hc.size()) < ns) {
        const int b = int(hc.size());
        hc.emplace_back(size_t(b + 1));
        for (int a = 0; a <= b; ++a) hc[b][a] = dot(basis[a], ab[b]);
      }
      std::vector<double> h(size_t(ns) * ns);
      for (int b = 0; b < ns; ++b)
        for (int a = 0; a <= b; ++a) h[a * ns + b] = h[b * ns + a] = hc[b][a];
      std::vector<double> ev, vv;
      jacobi_small(h, ns, ev, vv);
      int mi = int(std::min_element(ev.begin(), ev.end()) - ev.begin());
      eig = ev[mi];
      std::vector<double> ay(static_cast<size_t>(dim), 0.0);
      std::fill(y.begin(), y.end(), 0.0);
      for (int a = 0; a < ns; ++a) {
        const double c = vv[a * ns + mi];
        for (long j = 0; j < dim; ++j) {
          y[j] += c * basis[a][j];
          ay[j] += c * ab[a][j];
        }
      }
      std::vector<double> r(static_cast<size_t>(dim));
      for (long j = 0; j < dim; ++j) r[j] = ay[j] - eig * y[j];
      const double rn = std::sqrt(dot(r, r));
      if (rn < tol) break;
      std::vector<double> t = r;
      if (diag)
        for (long j = 0; j < dim; ++j) {
          const double den = eig - (*diag)[j];
          if (std::fabs(den) > 1e-14) t[j] /= den;
        }
      auto orth = [&](std::vector<double>& u) {
        for (int pass = 0; pass < 2; ++pass)
          for (auto& bv : basis) {
            const double d = dot(bv, u);
            for (long j = 0; j < dim; ++j) u[j] -= d * bv[j];
          }
        return std::sqrt(dot(u, u));
      };
      double cn = orth(t);
      if (cn < 1e-14) {
        t = r;
        cn = orth(t);
        if (cn < 1e-14) break;
      }
      for (double& e : t) e /= cn;
      if (ns >= max_sub) {
        if (ns >= dim) break;
        basis.assign(1, y);
        ab.assign(1, ay);
        hc.clear();
      }
      basis.push_back(t);
      ab.emplace_back();
      matvec(basis.back(), ab.back());
    }
    nrm = std::sqrt(dot(y, y));
    for (long j = 0; j < dim; ++j) x[j] = y[j] / nrm;
    return eig;
}

}  // namespace

// truncate_svd's selection (truncation/mod.rs:56-98) over the singular values
// of several blocks (each run descending): candidates by sigma descending
// (ties by block, then position); keep while count < max_bond and sigma >=
// cutoff * sigma_max, treating sigma^2 < floor * sigma_max^2 as zero; with
// deg_tol > 0 never cut between two values closer than deg_tol * sigma_max
// (the cut moves up to the group's start unless nothing would remain); at
// least min(min_bond, candidates).  Returns per-block kept counts, the kept
// and discarded totals and the summed sigma^2 of all and of the kept values.
void select_keep(const std::vector<std::vector<double>>& sig, long max_bond, double cutoff, long min_bond,
                 double deg_tol, double floor, std::vector<long>& keep_of, long* n_kept, long* n_disc,
                 double* sum_all, double* sum_kept) {
  struct Cand { double l; int mid; int idx; };
  std::vector<Cand> all;
  for (int k = 0; k < int(sig.size()); ++k)
    for (int j = 0; j < int(sig[k].size()); ++j) all.push_back({sig[k][j] * sig[k][j], k, j});
  std::stable_sort(all.begin(), all.end(), [](const Cand& a, const Cand& b) {
    if (a.l != b.l) return a.l > b.l;
    if (a.mid != b.mid) return a.mid < b.mid;
    return a.idx < b.idx;
  });
  keep_of.assign(sig.size(), 0);
  const double lmax = all.empty() ? 0 : all[0].l;
  const double thr = std::max(cutoff * cutoff, floor) * lmax;
  size_t keep = 0;
  while (keep < all.size() && long(keep) < max_bond && all[keep].l >= thr) ++keep;
  if (deg_tol > 0 && keep < all.size() && keep > 0) {
    size_t k2 = keep;
    const double smax = std::sqrt(lmax);
    while (k2 > 0 && std::sqrt(all[k2 - 1].l) - std::sqrt(all[k2].l) <= deg_tol * smax) --k2;
    if (k2 > 0) keep = k2;
  }
  keep = std::max<size_t>(keep, size_t(std::min<long>(std::max<long>(min_bond, 0), long(all.size()))));
  double sa = 0, sk = 0;
  for (size_t k = 0; k < all.size(); ++k) {
    sa += all[k].l;
    if (k < keep) {
      sk += all[k].l;
      ++keep_of[all[k].mid];
    }
  }
  *n_kept = long(keep);
  *n_disc = long(all.size() - keep);
  *sum_all = sa;
  *sum_kept = sk;
}

struct DmrgCfg {
  int sym = 1;
  std::vector<long> schedule;
  double cutoff = 1e-10, deg_tol = 1e-12, tol = 1e-10;
  long min_bond = 1;
  int max_sub = 60, max_iter = 1000;
  bool precondition = true;
  int nthreads = 0;
};

struct Dmrg {
  DmrgCfg cfg;
  int n = 0;
  std::vector<Mpo> mpo;
  std::vector<Leg> bonds;
  std::vector<Tensor> A;
  std::vector<Tensor> L, R;
  double energy = 0;
  int sweeps = 0;
  // per-sweep / per-step stats
  double max_trunc = 0;
  long max_bond = 0;
  int matvecs = 0;
  double t_matvec = 0, t_svd = 0, t_env = 0;

  Tensor site_tensor(int i) const {
    return make_tensor(cfg.sym, {mpo[i].phys, bonds[i], bonds[i + 1]}, {1, 1, 0}, Q{0, 0});
  }

  Tensor theta(int i) const {
    Tensor th = make_psi2(cfg.sym, mpo[i].phys, mpo[i + 1].phys, bonds[i], bonds[i + 2]);
    const Leg &Pi = mpo[i].phys, &Pj = mpo[i + 1].phys, &BL = bonds[i], &BM = bonds[i + 1], &BR = bonds[i + 2];
    for (auto& [k, b] : th.blocks) {
      const int si = b.sec[0], sj = b.sec[1], ls = b.sec[2], rs = b.sec[3];
      const int ms = BM.find(add(BL.q[ls], Pi.q[si]));
      if (ms < 0) continue;
      const Block* a = A[i].get({si, ls, ms});
      const Block* c = A[i + 1].get({sj, ms, rs});
      if (!a || !c) continue;
      const long dL = BL.dim[ls], dR = BR.dim[rs], dm = BM.dim[ms];
      for (long xi = 0; xi < Pi.dim[si]; ++xi)
        for (long xj = 0; xj < Pj.dim[sj]; ++xj)
          gemm_nn(dL, dR, dm, 1.0, a->data.data() + xi * dL * dm, dm, c->data.data() + xj * dm * dR, dR,
                  b.data.data() + (xi * Pj.dim[sj] + xj) * dL * dR, dR);
    }
    return th;
  }

  std::vector<double> flat(const Tensor& t) const {
    std::vector<double> v(size_t(t.total()));
    t.store(v.data());
    return v;
  }

  // Davidson (eigensolver/mod.rs:325-516) with warm start and preconditioner.
  double davidson(int i, std::vector<double>& x, int* nmv) {
    Tensor xt = make_psi2(cfg.sym, mpo[i].phys, mpo[i + 1].phys, bonds[i], bonds[i + 2]);
    Tensor yt = xt;
    const long dim = xt.total();
    auto matvec = [&](const std::vector<double>& in, std::vector<double>& out) {
      double t0 = wall();
      xt.load(in.data());
      heff_apply(L[i], R[i + 2], mpo[i], mpo[i + 1], xt, yt, cfg.nthreads);
      out.resize(size_t(dim));
      yt.store(out.data());
      ++*nmv;
      t_matvec += wall() - t0;
    };
    std::vector<double> diag;
    if (cfg.precondition) {
      Tensor dt = xt;
      heff_diag(L[i], R[i + 2], mpo[i], mpo[i + 1], dt);
      diag = flat(dt);
    }
    return davidson_core(matvec, dim, cfg.precondition ? &diag : nullptr, cfg.max_sub, cfg.max_iter, cfg.tol, x);
  }

  // SVD of the two-site vector per middle charge, truncation, write-back.
  void split(int i, bool lr, const Tensor& th, long max_bond_dim, double cutoff, double floor, double* terr,
             long* bdim) {
    double t0 = wall();
    const Leg &Pi = mpo[i].phys, &Pj = mpo[i + 1].phys, &BL = bonds[i], &BR = bonds[i + 2];
    struct Part { int s, x, b; long off, d; };
    struct Mid { Q q; std::vector<Part> rows, cols; long nr = 0, nc = 0; std::vector<double> U, S, V; long keep = 0; };
    std::map<Q, Mid> mm;
    for (int s = 0; s < Pi.n(); ++s)
      for (int x = 0; x < Pi.dim[s]; ++x)
        for (int l = 0; l < BL.n(); ++l) {
          Mid& m = mm[add(Pi.q[s], BL.q[l])];
          m.rows.push_back({s, x, l, m.nr, BL.dim[l]});
          m.nr += BL.dim[l];
        }
    for (int s = 0; s < Pj.n(); ++s)
      for (int x = 0; x < Pj.dim[s]; ++x)
        for (int r = 0; r < BR.n(); ++r) {
          Mid& m = mm[add(BR.q[r], neg(Pj.q[s]))];
          m.cols.push_back({s, x, r, m.nc, BR.dim[r]});
          m.nc += BR.dim[r];
        }
    std::vector<Mid> mids;
    for (auto& [q, m] : mm)
      if (m.nr > 0 && m.nc > 0) {
        m.q = q;
        mids.push_back(std::move(m));
      }
    auto slice = [&](const Part& rp, const Part& cp) -> const double* {
      const Block* b = th.get({rp.s, cp.s, rp.b, cp.b});
      if (!b) throw std::runtime_error("missing theta block");
      return b->data.data() + (rp.x * b->dims[1] + cp.x) * rp.d * cp.d;
    };
#pragma omp parallel for schedule(dynamic, 1) num_threads(cfg.nthreads > 0 ? cfg.nthreads : omp_get_max_threads())
    for (long k = 0; k < long(mids.size()); ++k) {
      Mid& m = mids[k];
      std::vector<double> M(size_t(m.nr * m.nc));
      for (auto& rp : m.rows)
        for (auto& cp : m.cols) {
          const double* src = slice(rp, cp);
          for (long a = 0; a < rp.d; ++a)
            for (long b = 0; b < cp.d; ++b) M[size_t((rp.off + a) * m.nc + cp.off + b)] = src[a * cp.d + b];
        }
      svd_jacobi(M, m.nr, m.nc, m.U, m.S, m.V);
    }
    std::vector<std::vector<double>> sv(mids.size());
    for (size_t k = 0; k < mids.size(); ++k) sv[k] = mids[k].S;
    std::vector<long> kb;
    long nk = 0, nd = 0;
    double sa = 0, sk = 0;
    select_keep(sv, max_bond_dim, cutoff, cfg.min_bond, cfg.deg_tol, floor, kb, &nk, &nd, &sa, &sk);
    for (size_t k = 0; k < mids.size(); ++k) mids[k].keep = kb[k];
    *terr = sa > 0 ? (sa - sk) / sa : 0.0;
    const double rn = sk > 0 ? 1.0 / std::sqrt(sk) : 1.0;
    std::vector<std::pair<Q, long>> nb;
    std::map<Q, int> mid_of;
    for (int k = 0; k < int(mids.size()); ++k)
      if (mids[k].keep > 0) {
        nb.push_back({mids[k].q, mids[k].keep});
        mid_of[mids[k].q] = k;
      }
    Leg newB = make_leg_v(nb);
    *bdim = newB.total;
    Tensor Ai = make_tensor(cfg.sym, {Pi, BL, newB}, {1, 1, 0}, Q{0, 0});
    Tensor Aj = make_tensor(cfg.sym, {Pj, newB, BR}, {1, 1, 0}, Q{0, 0});
    for (auto& [key, b] : Ai.blocks) {
      const Mid& m = mids[mid_of.at(newB.q[b.sec[2]])];
      const long dL = BL.dim[b.sec[1]], kk = m.keep, r = long(m.S.size());
      for (long x = 0; x < Pi.dim[b.sec[0]]; ++x) {
        const Part* rp = nullptr;
        for (auto& p : m.rows)
          if (p.s == b.sec[0] && p.x == x) rp = &p;
        for (long a = 0; a < dL; ++a)
          for (long c = 0; c < kk; ++c) {
            double u = m.U[size_t((rp->off + a) * r + c)];
            if (!lr) u *= m.S[c] * rn;  // centre U S on the left in a right-to-left step
            b.data[size_t((x * dL + a) * kk + c)] = u;
          }
      }
    }
    for (auto& [key, b] : Aj.blocks) {
      const Mid& m = mids[mid_of.at(newB.q[b.sec[1]])];
      const long dR = BR.dim[b.sec[2]], kk = m.keep, r = long(m.S.size());
      for (long x = 0; x < Pj.dim[b.sec[0]]; ++x) {
        const Part* cp = nullptr;
        for (auto& p : m.cols)
          if (p.s == b.sec[0] && p.x == x) cp = &p;
        for (long c = 0; c < kk; ++c)
          for (long a = 0; a < dR; ++a) {
            double v = m.V[size_t((cp->off + a) * r + c)];
            if (lr) v *= m.S[c] * rn;  // centre S V^T on the right in a left-to-right step
            b.data[size_t((x * kk + c) * dR + a)] = v;
          }
      }
    }
    bonds[i + 1] = newB;
    A[i] = std::move(Ai);
    A[i + 1] = std::move(Aj);
    t_svd += wall() - t0;
  }

  // L[i+1] from L[i] and A[i]
  void grow_left(int i) {
    double t0 = wall();
    const Mpo& W = mpo[i];
    const Leg &P = W.phys, &BL = bonds[i], &BR = bonds[i + 1];
    Tensor out = make_env(cfg.sym, BR, W.wr);
    std::vector<Block*> ob;
    for (auto& [k, b] : out.blocks) ob.push_back(&b);
    const Tensor& Lo = L[i];
    const Tensor& Ai = A[i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(cfg.nthreads > 0 ? cfg.nthreads : omp_get_max_threads())
    for (long e = 0; e < long(ob.size()); ++e) {
      Block& o = *ob[e];
      const int rps = o.sec[0], swr = o.sec[1], rs = o.sec[2];
      const long dRp = BR.dim[rps], dR = BR.dim[rs], dwr = W.wr.dim[swr];
      std::vector<double> tmp;
      for (long so = 0; so < P.total; ++so)
        for (long s = 0; s < P.total; ++s)
          for (long wl = 0; wl < W.wl.total; ++wl)
            for (long xwr = 0; xwr < dwr; ++xwr) {
              const long wr = W.wr.off[swr] + xwr;
              const double c = W.at(s, so, wl, wr);
              if (c == 0.0) continue;
              int sso, ss, swl;
              long xso, xs, xwl;
              locate(P, so, &sso, &xso);
              locate(P, s, &ss, &xs);
              locate(W.wl, wl, &swl, &xwl);
              const int lps = BL.find(add(BR.q[rps], neg(P.q[sso])));
              if (lps < 0) continue;
              const int ls = BL.find(add(BL.q[lps], W.wl.q[swl]));
              if (ls < 0) continue;
              const Block* bra = Ai.get({sso, lps, rps});
              const Block* ket = Ai.get({ss, ls, rs});
              const Block* lb = Lo.get({lps, swl, ls});
              if (!bra || !ket || !lb) continue;
              const long dLp = BL.dim[lps], dL = BL.dim[ls], dwl = W.wl.dim[swl];
              // tmp (dLp x dR) = L_{wl}(a'_L, a_L) A[s](a_L, a_R)
              tmp.assign(size_t(dLp * dR), 0.0);
              gemm_nn(dLp, dR, dL, 1.0, lb->data.data() + xwl * dL, dwl * dL, ket->data.data() + xs * dL * dR, dR,
                      tmp.data(), dR);
              // out(a'_R, w_r, a_R) += c * A[s'](a'_L, a'_R)^T tmp
              gemm_tn(dRp, dR, dLp, c, bra->data.data() + xso * dLp * dRp, dRp, tmp.data(), dR,
                      o.data.data() + xwr * dR, dwr * dR);
            }
    }
    L[i + 1] = std::move(out);
    t_env += wall() - t0;
  }

  // R[i] from R[i+1] and A[i]
  void grow_right(int i) {
    double t0 = wall();
    const Mpo& W = mpo[i];
    const Leg &P = W.phys, &BL = bonds[i], &BR = bonds[i + 1];
    Tensor out = make_env(cfg.sym, BL, W.wl);
    std::vector<Block*> ob;
    for (auto& [k, b] : out.blocks) ob.push_back(&b);
    const Tensor& Ro = R[i + 1];
    const Tensor& Ai = A[i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(cfg.nthreads > 0 ? cfg.nthreads : omp_get_max_threads())
    for (long e = 0; e < long(ob.size()); ++e) {
      Block& o = *ob[e];
      const int lps = o.sec[0], swl = o.sec[1], ls = o.sec[2];
      const long dLp = BL.dim[lps], dL = BL.dim[ls], dwl = W.wl.dim[swl];
      std::vector<double> tmp;
      for (long so = 0; so < P.total; ++so)
        for (long s = 0; s < P.total; ++s)
          for (long xwl = 0; xwl < dwl; ++xwl)
            for (long wr = 0; wr < W.wr.total; ++wr) {
              const long wl = W.wl.off[swl] + xwl;
              const double c = W.at(s, so, wl, wr);
              if (c == 0.0) continue;
              int sso, ss, swr;
              long xso, xs, xwr;
              locate(P, so, &sso, &xso);
              locate(P, s, &ss, &xs);
              locate(W.wr, wr, &swr, &xwr);
              const int rps = BR.find(add(BL.q[lps], P.q[sso]));
              if (rps < 0) continue;
              const int rs = BR.find(add(BR.q[rps], W.wr.q[swr]));
              if (rs < 0) continue;
              const Block* bra = Ai.get({sso, lps, rps});
              const Block* ket = Ai.get({ss, ls, rs});
              const Block* rb = Ro.get({rps, swr, rs});
              if (!bra || !ket || !rb) continue;
              const long dRp = BR.dim[rps], dR = BR.dim[rs], dwr = W.wr.dim[swr];
              // tmp (dLp x dR) = A[s'](a'_L, a'_R) R_{wr}(a'_R, a_R)
              tmp.assign(size_t(dLp * dR), 0.0);
              gemm_nn(dLp, dR, dRp, 1.0, bra->data.data() + xso * dLp * dRp, dRp, rb->data.data() + xwr * dR,
                      dwr * dR, tmp.data(), dR);
              // out(a'_L, w_l, a_L) += c * tmp A[s](a_L, a_R)^T
              gemm_nt(dLp, dL, dR, c, tmp.data(), dR, ket->data.data() + xs * dL * dR, dR,
                      o.data.data() + xwl * dL, dwl * dL);
            }
    }
    R[i] = std::move(out);
    t_env += wall() - t0;
  }

  Tensor boundary(const Leg& b, const Leg& w) {
    Tensor t = make_env(cfg.sym, b, w);
    for (auto& [k, blk] : t.blocks) blk.data[0] = 1.0;
    return t;
  }

  void prepare() {
    for (int i = n - 2; i >= 0; --i) {
      Tensor th = theta(i);
      double te;
      long bd;
      split(i, false, th, 1L << 40, 0.0, kLambdaZero, &te, &bd);
    }
    L.assign(n + 1, Tensor());
    R.assign(n + 1, Tensor());
    L[0] = boundary(bonds[0], mpo[0].wl);
    R[n] = boundary(bonds[n], mpo[n - 1].wr);
    for (int i = n - 1; i >= 2; --i) grow_right(i);
  }

  double step(int i, bool lr, long D, double* terr, long* bdim, int* nmv) {
    Tensor th = theta(i);
    std::vector<double> x = flat(th);
    const double e = davidson(i, x, nmv);
    th.load(x.data());
    split(i, lr, th, D, cfg.cutoff, std::max(cfg.cutoff * cfg.cutoff, kLambdaZero), terr, bdim);
    if (lr) grow_left(i);
    else grow_right(i + 1);
    return e;
  }

  double sweep(double* seconds) {
    const long D = cfg.schedule[std::min<size_t>(size_t(sweeps), cfg.schedule.size() - 1)];
    double t0 = wall();
    max_trunc = 0;
    max_bond = 0;
    matvecs = 0;
    for (int i = 0; i + 1 < n; ++i) {
      double te;
      long bd;
      energy = step(i, true, D, &te, &bd, &matvecs);
      max_trunc = std::max(max_trunc, te);
      max_bond = std::max(max_bond, bd);
    }
    for (int i = n - 2; i >= 1; --i) {
      double te;
      long bd;
      energy = step(i, false, D, &te, &bd, &matvecs);
      max_trunc = std::max(max_trunc, te);
      max_bond = std::max(max_bond, bd);
    }
    ++sweeps;
    *seconds = wall() - t0;
    return energy;
  }
};

}  // namespace orc

using namespace orc;

extern "C" {

void* or_dmrg_create(int32_t sym, int32_t n, const or_qindex* phys, const or_qindex* wl, const or_qindex* wr,
                     const double* const* w, const int64_t* schedule, int32_t n_schedule, double cutoff,
                     double deg_tol, double tol, int32_t max_sub, int32_t max_iter, int32_t precondition,
                     int32_t nthreads) {
  try {
    auto d = std::make_unique<Dmrg>();
    d->n = n;
    d->cfg.sym = sym;
    d->cfg.schedule.assign(schedule, schedule + n_schedule);
    d->cfg.cutoff = cutoff;
    d->cfg.deg_tol = deg_tol;
    d->cfg.tol = tol;
    d->cfg.max_sub = max_sub;
    d->cfg.max_iter = max_iter;
    d->cfg.precondition = precondition != 0;
    d->cfg.nthreads = nthreads;
    for (int i = 0; i < n; ++i)
      d->mpo.push_back(make_mpo(make_leg(phys[i]), make_leg(wl[i]), make_leg(wr[i]), w[i]));
    return d.release();
  } catch (...) {
    return nullptr;
  }
}

void or_dmrg_destroy(void* h) { delete static_cast<Dmrg*>(h); }

int or_dmrg_set_state(void* h, const or_qindex* bonds, const double* const* data) {
  try {
    Dmrg* d = static_cast<Dmrg*>(h);
    d->bonds.clear();
    for (int i = 0; i <= d->n; ++i) d->bonds.push_back(make_leg(bonds[i]));
    d->A.clear();
    for (int i = 0; i < d->n; ++i) {
      d->A.push_back(d->site_tensor(i));
      d->A.back().load(data[i]);
    }
    d->sweeps = 0;
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_dmrg_prepare(void* h) {
  try {
    static_cast<Dmrg*>(h)->prepare();
    return 0;
  } catch (...) {
    return 1;
  }
}

// out: energy, max truncation error, max bond dim, seconds, matvecs, t_matvec, t_svd, t_env
int or_dmrg_sweep(void* h, double* out) {
  try {
    Dmrg* d = static_cast<Dmrg*>(h);
    d->t_matvec = d->t_svd = d->t_env = 0;
    double secs;
    const double e = d->sweep(&secs);
    out[0] = e;
    out[1] = d->max_trunc;
    out[2] = double(d->max_bond);
    out[3] = secs;
    out[4] = double(d->matvecs);
    out[5] = d->t_matvec;
    out[6] = d->t_svd;
    out[7] = d->t_env;
    return 0;
  } catch (...) {
    return 1;
  }
}

// Dense-matrix entry points of the solver pieces, for the known-answer tests.
int or_jacobi_small(int32_t n, const double* h, double* evals, double* evecs) {
  std::vector<double> m(h, h + size_t(n) * n), ev, vv;
  jacobi_small(m, n, ev, vv);
  std::copy(ev.begin(), ev.end(), evals);
  std::copy(vv.begin(), vv.end(), evecs);
  return 0;
}

// x0 == NULL: the reference start vector ((7i+13) mod 97)/97 - 0.5.
int or_davidson_dense(int32_t n, const double* h, const double* diag, const double* x0, double tol, int32_t max_sub,
                      double* eigval, double* x_out, int32_t* n_matvec) {
  std::vector<double> x(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) x[i] = x0 ? x0[i] : double((i * 7 + 13) % 97) / 97.0 - 0.5;
  int nmv = 0;
  auto mv = [&](const std::vector<double>& in, std::vector<double>& out) {
    out.assign(size_t(n), 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) out[i] += h[size_t(i) * n + j] * in[j];
    ++nmv;
  };
  std::vector<double> d;
  if (diag) d.assign(diag, diag + n);
  *eigval = davidson_core(mv, n, diag ? &d : nullptr, max_sub, 1000, tol, x);
  if (x_out) std::copy(x.begin(), x.end(), x_out);
  if (n_matvec) *n_matvec = nmv;
  return 0;
}

int or_truncation_select(int32_t n_blocks, const int64_t* counts, const double* sigma, int64_t max_bond,
                         double cutoff, int64_t min_bond, double deg_tol, double floor, int64_t* keep,
                         int64_t* n_kept, int64_t* n_disc, double* terr) {
  try {
    std::vector<std::vector<double>> sig(static_cast<size_t>(n_blocks));
    long at = 0;
    for (int b = 0; b < n_blocks; ++b)
      for (long j = 0; j < counts[b]; ++j) sig[size_t(b)].push_back(sigma[at++]);
    std::vector<long> kb;
    long nk = 0, nd = 0;
    double sa = 0, sk = 0;
    orc::select_keep(sig, max_bond, cutoff, min_bond, deg_tol, floor, kb, &nk, &nd, &sa, &sk);
    for (int b = 0; b < n_blocks; ++b) keep[b] = kb[size_t(b)];
    *n_kept = nk;
    *n_disc = nd;
    *terr = sa > 0 ? (sa - sk) / sa : 0.0;
    return 0;
  } catch (...) {
    return 1;
  }
}

int or_svd(int64_t m, int64_t n, const double* a, double* u, double* s, double* v) {
  std::vector<double> A(a, a + m * n), U, S, V;
  svd_jacobi(A, m, n, U, S, V);
  std::copy(U.begin(), U.end(), u);
  std::copy(S.begin(), S.end(), s);
  std::copy(V.begin(), V.end(), v);
  return 0;
}

int or_dmrg_bond(void* h, int32_t bond, int32_t capacity, int32_t* n_sectors, int32_t* qn, int64_t* dims) {
  Dmrg* d = static_cast<Dmrg*>(h);
  if (bond < 0 || bond > d->n) return 1;
  const Leg& l = d->bonds[bond];
  *n_sectors = l.n();
  if (l.n() > capacity) return 7;
  for (int i = 0; i < l.n(); ++i) {
    qn[2 * i] = l.q[i].a;
    qn[2 * i + 1] = l.q[i].b;
    dims[i] = l.dim[i];
  }
  return 0;
}

}  // extern "C"
