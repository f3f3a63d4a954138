// Davidson eigensolver with device-resident FP64 Krylov vectors (see davidson.hpp).
#include "davidson.hpp"

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "tdeig.hpp"

namespace tk {
namespace {

__global__ void residual_dev_kernel(const double* __restrict__ ay, const double* __restrict__ y,
                                    const double* __restrict__ theta, double* __restrict__ r, int64_t n) {
  const double th = *theta;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    r[j] = ay[j] - th * y[j];
}
__global__ void precond_dev_kernel(const double* __restrict__ r, const double* __restrict__ d,
                                   const double* __restrict__ theta, double* __restrict__ t, int64_t n) {
  const double th = *theta;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x) {
    const double den = th - d[j];
    t[j] = fabs(den) > 1e-14 ? r[j] / den : r[j];
  }
}
__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += int64_t(gridDim.x) * blockDim.x)
    b[j] = a[j];
}

int grid1(Context* ctx, int64_t n) {
  return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, int64_t(ctx->n_cu) * 16)));
}

struct Workspace {
  DBuf<double> V, AV, H, c, h, scal, y, ay, r;
  DBuf<int> flag;
};

}  // namespace

DavidsonResult davidson_generic(Context* ctx, int64_t n, MatvecFn mv, void* user, const double* diag_dev,
                                bool low_precision, double* x, const DavidsonOpts& opts) {
  DavidsonResult res;
  if (n <= 0) throw ArgError("davidson: empty problem");
  const int m = int(std::max<int64_t>(1, std::min<int64_t>({int64_t(opts.max_subspace), n, 64})));
  hipStream_t s = ctx->stream;
  Workspace w;
  w.V.reset(ctx, size_t(m) * n);
  w.AV.reset(ctx, size_t(m) * n);
  w.H.reset(ctx, size_t(m) * m);
  w.c.reset(ctx, m);
  w.h.reset(ctx, 2 * (size_t(m) + 1));
  w.scal.reset(ctx, 8);
  w.y.reset(ctx, n);
  w.ay.reset(ctx, n);
  w.r.reset(ctx, n);
  w.flag.reset(ctx, 1);
  double* V = w.V.get();
  double* AV = w.AV.get();
  double* H = w.H.get();
  double* c = w.c.get();
  double* theta = w.scal.get();
  double* rn2 = w.scal.get() + 1;
  double* tn2 = w.scal.get() + 2;
  const double* dprec = (opts.precondition && diag_dev) ? diag_dev : nullptr;
  // the first matvec and residual are FP64; low precision takes over only
  // for a cold start (see DavidsonOpts::low_start_rel)
  bool low = false;
  // corrections of a converging solve in low precision (DavidsonOpts::low_tail)
  bool tail = false, tail_done = false;

  auto matvec = [&](const double* in, double* out) {
    const bool lp = low || tail;
    mv(in, out, lp ? 0 : 1, user);
    ++res.matvecs;
    if (!lp) ++res.matvecs_fp64;
  };
  // basis <- {v / |v|}, AV0 = A v0, H00
  auto start_from = [&](const double* v) {
    if (v != V) hipLaunchKernelGGL(copy_kernel, dim3(grid1(ctx, n)), dim3(256), 0, s, v, V, n);
    dot_many(ctx, V, n, 1, V, n, tn2);
    scale_by_inv_sqrt(ctx, V, tn2, n);
    matvec(V, AV);
    dot_many(ctx, V, n, 1, AV, n, H);
  };
  // slot <- slot minus its projection on V[0..nsub): one classical pass, and
  // a second one (flag on the device) when the first removed more than half
  // of its squared norm; then *tn2 = |slot|^2
  auto orthogonalise = [&](double* slot, int nsub) {
    dot_many(ctx, V, n, nsub + 1, slot, n, w.h.get());  // V^T slot and slot.slot (slot = V[nsub])
    subtract_projection(ctx, V, n, nsub, w.h.get(), slot, n);
    reorth_flag(ctx, w.h.get(), nsub, w.flag.get());
    dot_many(ctx, V, n, nsub, slot, n, w.h.get(), w.flag.get());
    subtract_projection(ctx, V, n, nsub, w.h.get(), slot, n, w.flag.get());
    dot_many(ctx, slot, n, 1, slot, n, tn2);
  };
  // the same with V^T slot and slot.slot in h already (summed by the
  // residual pass): the first pass also sums V^T of its result, which the
  // second pass (when flagged) subtracts; each pass leaves |slot|^2 in *tn2
  auto project_rest = [&](double* slot, int nsub) {
    double* h2 = w.h.get() + (m + 1);
    project_out_dots(ctx, V, n, nsub, w.h.get(), slot, n, h2, tn2);
    reorth_flag(ctx, w.h.get(), nsub, w.flag.get());
    project_out_dots(ctx, V, n, nsub, h2, slot, n, nullptr, tn2, w.flag.get());
  };
  // the Ritz residual (preconditioned by pre unless null) into t and its norm
  // into *rn2; with `grow`, t is the new basis slot and is orthogonalised
  // against V (for a small basis in two passes over V fewer)
  auto residual_into = [&](double* t, int nsub, const double* pre, bool grow) {
    if (grow && nsub <= fused_vec_max()) {
      ritz_residual_dots(ctx, V, AV, n, nsub, c, theta, pre, t, n, rn2, w.h.get());
      project_rest(t, nsub);
    } else {
      ritz_residual(ctx, V, AV, n, nsub, c, theta, pre, t, n, rn2);
      if (grow) orthogonalise(t, nsub);
    }
  };
  // Ritz vector y = V c of the current projected solution
  auto ritz_vector = [&](int nsub) { combine(ctx, V, n, nsub, c, 0.0, w.y.get(), n); };
  start_from(x);
  int nsub = 1;
  double best = INFINITY;
  int stall = 0;
  double host[3] = {0, 0, 0};
  bool c_current = false;
  for (int it = 0; it < opts.max_iter; ++it) {
    // lowest Ritz pair of the projected matrix (one wavefront on the device)
    lowest_eigpair(ctx, H, m, nsub, c, theta);
    c_current = true;
    const bool can_grow = nsub < n;
    if (can_grow && nsub == m) {
      // restart from the Ritz vector; A y is carried over
      ritz_vector(nsub);
      combine(ctx, AV, n, nsub, c, 0.0, w.ay.get(), n);
      hipLaunchKernelGGL(copy_kernel, dim3(grid1(ctx, n)), dim3(256), 0, s, w.y.get(), V, n);
      hipLaunchKernelGGL(copy_kernel, dim3(grid1(ctx, n)), dim3(256), 0, s, w.ay.get(), AV, n);
      dot_many(ctx, V, n, 1, AV, n, H);
      fill_f64(ctx, c, 1.0, 1);
      nsub = 1;
    }
    // the residual, its norm and the next correction vector in one pass over
    // V and AV, formed before the one host sync of the iteration (discarded
    // when the iteration converges or switches precision)
    double* slot = V + int64_t(nsub) * n;
    residual_into(can_grow ? slot : w.r.get(), nsub, dprec, can_grow);
    TK_HIP(hipMemcpyAsync(host, w.scal.get(), 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    TK_HIP(hipGetLastError());
    res.iterations = it + 1;
    res.eigenvalue = host[0];
    res.residual = std::sqrt(std::max(0.0, host[1]));
    if (it == 0 && low_precision &&
        res.residual > opts.low_start_rel * std::max(1.0, std::fabs(res.eigenvalue)))
      low = true;  // far from converged: expand the basis with low-precision matvecs
    if (tail) {
      // a residual that stops falling has met the rounding of the
      // low-precision products: refine in FP64 from the Ritz vector
      if (res.residual < best * 0.7) {
        best = res.residual;
        stall = 0;
      } else if (++stall >= 3 && res.residual >= opts.tol) {
        tail = false;
        tail_done = true;
        ritz_vector(nsub);
        start_from(w.y.get());
        fill_f64(ctx, c, 1.0, 1);
        nsub = 1;
        continue;
      }
    }
    if (low) {
      if (res.residual < best * 0.7) {
        best = res.residual;
        stall = 0;
      } else {
        ++stall;
      }
      if (res.residual < opts.switch_rel * std::max(1.0, std::fabs(res.eigenvalue)) || stall >= 3 ||
          res.residual < opts.tol) {
        low = false;  // FP64 residual refinement from the current Ritz vector
        ritz_vector(nsub);
        start_from(w.y.get());
        fill_f64(ctx, c, 1.0, 1);
        nsub = 1;
        continue;
      }
    } else if (res.residual < opts.tol) {
      res.converged = true;
      break;
    } else if (low_precision && !tail && !tail_done && opts.low_tail > 0 && res.residual < opts.low_tail * opts.tol) {
      tail = true;
      best = res.residual;
      stall = 0;
    }
    if (!can_grow) break;  // the basis spans the space: exact
    if (std::sqrt(std::max(0.0, host[2])) < 1e-14) {
      // the preconditioned correction collapsed into the basis: raw residual
      residual_into(slot, nsub, nullptr, true);
      TK_HIP(hipMemcpyAsync(host + 2, tn2, sizeof(double), hipMemcpyDeviceToHost, s));
      TK_HIP(hipStreamSynchronize(s));
      if (std::sqrt(std::max(0.0, host[2])) < 1e-14) {
        if (low || tail) {  // stagnated in low precision: refine in FP64
          low = false;
          if (tail) tail_done = true;
          tail = false;
          ritz_vector(nsub);
          start_from(w.y.get());
          fill_f64(ctx, c, 1.0, 1);
          nsub = 1;
          continue;
        }
        break;
      }
    }
    scale_by_inv_sqrt(ctx, slot, tn2, n);
    matvec(slot, AV + int64_t(nsub) * n);
    if (tail) {
      // H_jt = (A v_j) . t from the products held (A is symmetric), so the
      // couplings of the new vector to the FP64 columns carry no
      // low-precision rounding; only H_tt uses the new product
      dot_many(ctx, AV, n, nsub, slot, n, H + int64_t(nsub) * m);
      dot_many(ctx, slot, n, 1, AV + int64_t(nsub) * n, n, H + int64_t(nsub) * m + nsub);
    } else {
      dot_many(ctx, V, n, nsub + 1, AV + int64_t(nsub) * n, n, H + int64_t(nsub) * m);
    }
    ++nsub;
    c_current = false;
  }
  // eigenvector: normalised Ritz vector (of the grown basis when the
  // iteration budget ran out after a matvec)
  if (!c_current) lowest_eigpair(ctx, H, m, nsub, c, theta);
  ritz_vector(nsub);
  dot_many(ctx, w.y.get(), n, 1, w.y.get(), n, tn2);
  scale_by_inv_sqrt(ctx, w.y.get(), tn2, n);
  hipLaunchKernelGGL(copy_kernel, dim3(grid1(ctx, n)), dim3(256), 0, s, w.y.get(), x, n);
  TK_HIP(hipGetLastError());
  ctx->sync();
  return res;
}

namespace {
__global__ void sqrt_to_kernel(const double* __restrict__ nrm2, double* __restrict__ out) {
  if (threadIdx.x == 0) *out = sqrt(fmax(0.0, *nrm2));
}
}  // namespace

DavidsonResult lanczos_generic(Context* ctx, int64_t n, MatvecFn mv, void* user, bool low_precision, double* x,
                               const LanczosOpts& opts) {
  DavidsonResult res;
  if (n <= 0) throw ArgError("lanczos: empty problem");
  const int m = int(std::max<int64_t>(1, std::min<int64_t>({int64_t(opts.max_krylov), n, 512})));
  hipStream_t s = ctx->stream;
  DBuf<double> Vb(ctx, size_t(m) * n), wb(ctx, n), yb(ctx, n), ab(ctx, m), bb(ctx, m), hb(ctx, m), zb(ctx, m),
      sc(ctx, 8);
  double* V = Vb.get();
  double* w = wb.get();
  double* y = yb.get();
  double* alpha = ab.get();
  double* beta = bb.get();
  double* theta = sc.get();
  double* est = sc.get() + 1;
  double* nrm2 = sc.get() + 2;
  bool low = low_precision;
  auto matvec = [&](const double* in, double* out) {
    mv(in, out, low ? 0 : 1, user);
    ++res.matvecs;
    if (!low) ++res.matvecs_fp64;
  };
  auto normalise = [&](double* v) {
    dot_many(ctx, v, n, 1, v, n, nrm2);
    scale_by_inv_sqrt(ctx, v, nrm2, n);
  };
  TK_HIP(hipMemcpyAsync(y, x, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
  normalise(y);
  double host[3] = {0, 0, 0};
  for (int restart = 0; restart <= opts.max_restarts; ++restart) {
    TK_HIP(hipMemcpyAsync(V, y, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
    int kept = 0;
    bool stop = false, to_fp64 = false;
    for (int j = 0; j < m; ++j) {
      double* vj = V + int64_t(j) * n;
      matvec(vj, w);
      // alpha_j and full reorthogonalisation (two classical passes)
      dot_many(ctx, V, n, j + 1, w, n, hb.get());
      TK_HIP(hipMemcpyAsync(alpha + j, hb.get() + j, sizeof(double), hipMemcpyDeviceToDevice, s));
      subtract_projection(ctx, V, n, j + 1, hb.get(), w, n);
      dot_many(ctx, V, n, j + 1, w, n, hb.get());
      subtract_projection(ctx, V, n, j + 1, hb.get(), w, n);
      dot_many(ctx, w, n, 1, w, n, nrm2);
      hipLaunchKernelGGL(sqrt_to_kernel, dim3(1), dim3(64), 0, s, nrm2, beta + j);
      tridiag_lowest(ctx, alpha, beta, j + 1, beta + j, zb.get(), theta, est);
      TK_HIP(hipMemcpyAsync(host, sc.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
      TK_HIP(hipMemcpyAsync(host + 2, beta + j, sizeof(double), hipMemcpyDeviceToHost, s));
      TK_HIP(hipStreamSynchronize(s));
      TK_HIP(hipGetLastError());
      ++res.iterations;
      kept = j + 1;
      res.eigenvalue = host[0];
      res.residual = host[1];
      const double scale = std::max(1.0, std::fabs(host[0]));
      if (low && (host[1] < opts.switch_rel * scale || host[1] < opts.tol)) {
        to_fp64 = true;
        break;
      }
      if (!low && host[1] < opts.tol) {
        res.converged = true;
        stop = true;
        break;
      }
      if (host[2] < 1e-15 * scale) {  // invariant subspace: T holds exact eigenvalues
        if (low) {
          to_fp64 = true;
        } else {
          res.converged = true;
          stop = true;
        }
        break;
      }
      if (j + 1 < m) {
        double* vn = V + int64_t(j + 1) * n;
        TK_HIP(hipMemcpyAsync(vn, w, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
        scale_by_inv_sqrt(ctx, vn, nrm2, n);
      }
    }
    // Ritz vector y = V z
    combine(ctx, V, n, kept, zb.get(), 0.0, y, n);
    normalise(y);
    if (stop) break;
    if (to_fp64) low = false;
  }
  // explicit FP64 residual of the returned pair
  low = false;
  matvec(y, w);
  dot_many(ctx, y, n, 1, w, n, theta);
  hipLaunchKernelGGL(residual_dev_kernel, dim3(grid1(ctx, n)), dim3(256), 0, s, w, y, theta, V, n);
  dot_many(ctx, V, n, 1, V, n, nrm2);
  TK_HIP(hipMemcpyAsync(host, theta, sizeof(double), hipMemcpyDeviceToHost, s));
  TK_HIP(hipMemcpyAsync(host + 1, nrm2, sizeof(double), hipMemcpyDeviceToHost, s));
  TK_HIP(hipMemcpyAsync(x, y, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
  TK_HIP(hipStreamSynchronize(s));
  res.eigenvalue = host[0];
  res.residual = std::sqrt(std::max(0.0, host[1]));
  res.converged = res.converged && res.residual < std::max(opts.tol * 10.0, 1e-12);
  TK_HIP(hipGetLastError());
  return res;
}

namespace {
struct PlanUser {
  HeffPlan* plan;
};
// The plan's own precision replays its graph; FP64 products of a
// low-precision plan (the refinement) take the eager FP64 path.  With
// TK_HEFF_GRAPHS=0 in the environment every product is launched eagerly (the
// same kernels and work lists), for tools that cannot follow graph launches.
bool heff_graphs() {
  static const bool on = [] {
    const char* e = std::getenv("TK_HEFF_GRAPHS");
    return e == nullptr || std::atoi(e) != 0;
  }();
  return on;
}
void plan_matvec(const double* x, double* y, int fp64, void* user) {
  HeffPlan* p = static_cast<PlanUser*>(user)->plan;
  if (fp64 && p->prec() != Prec::FP64) p->apply_fp64(x, y);
  else if (heff_graphs()) p->apply_graph(x, y);
  else p->apply(x, y);
}
}  // namespace

DavidsonResult lanczos(Context* ctx, HeffPlan& plan, double* x, const LanczosOpts& opts) {
  PlanUser u{&plan};
  return lanczos_generic(ctx, plan.psi().total, plan_matvec, &u, plan.prec() != Prec::FP64, x, opts);
}

DavidsonResult davidson(Context* ctx, HeffPlan& plan, double* x, const DavidsonOpts& opts) {
  const int64_t n = plan.psi().total;
  DBuf<double> diag;
  if (opts.precondition) {
    diag.reset(ctx, n);
    plan.diagonal(diag.get());
  }
  PlanUser u{&plan};
  return davidson_generic(ctx, n, plan_matvec, &u, opts.precondition ? diag.get() : nullptr,
                          plan.prec() != Prec::FP64, x, opts);
}

}  // namespace tk
