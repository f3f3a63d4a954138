// Two-site DMRG on block-sparse tensors, every contraction on the GPU.
//
// Reference driver: DMRGEngine::run_with_cancel_flag / dmrg_step_two_site /
// SweepSchedule::standard / converged (tk-dmrg/src/sweep/mod.rs:59-81, 274-448,
// 517-526), in the spec'd step order (techspec/6_tech-spec_tk-dmrg.md:1033-1042),
// with the draft's defects fixed: the SVD factors are written back into the
// MPS, the right environment of bond (i, i+1) is R[i+2] (sites i+2..N-1), the
// two-site vector is matricised (s_i, a_L) x (s_j, a_R) before the SVD, and the
// eigensolver is warm-started from the current two-site tensor.
//
// Conventions
//   MPS site (s, a_L, a_R), dirs [Out, Out, In], flux 0: q(a_R) = q(s) + q(a_L)
//   bond 0 = {0:1}, bond N = {total charge:1}
//   L[i] on bond i covers sites 0..i-1; R[i] on bond i covers sites i..N-1
//   environments (bra, w, ket) with every allowed block, q(bra) = q(ket) - q(w)
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "davidson.hpp"
#include "device.hpp"
#include "heff.hpp"
#include "sym.hpp"

namespace tk {

// sigma^2 <= kLambdaZero sigma_max^2 counts as a numerical zero in every
// split (sigma <= 1e-13 sigma_max, below the reference's default cutoff).
constexpr double kLambdaZero = 1e-26;

struct DmrgConfig {
  Sym sym = Sym::U1;
  std::vector<MpoSite> mpo;              // one per site; mpo[i].wr == mpo[i+1].wl
  std::vector<int64_t> bond_schedule{200};  // max bond dimension per sweep (last value repeats)
  double svd_cutoff = 1e-12;             // relative: keep sigma >= cutoff * sigma_max
  int64_t min_bond_dim = 1;
  double degeneracy_tol = 1e-12;         // never cut inside a group |s_k - s_{k+1}| <= tol * s_max (s = sigma)
  Prec prec = Prec::FP64;
  int eigensolver = 0;  // 0 Davidson (DMRGConfig's default, sweep/mod.rs:169-177), 1 Lanczos
  DavidsonOpts davidson;
  LanczosOpts lanczos;
  double energy_tol = 1e-10;             // sweep convergence |dE| / max(|E|, 1)
};

struct StepStats {
  int site = 0;
  bool left_to_right = true;
  double energy = 0;
  double truncation_error = 0;  // 1 - sum kept sigma^2 / sum sigma^2
  int64_t bond_dim = 0;
  int iterations = 0, matvecs = 0, matvecs_fp64 = 0;
  bool converged = true;
  double residual = 0;
  int64_t dim = 0;        // two-site vector length
  int64_t svd_nmax = 0;   // largest Gram block of the split
  double svd_n3 = 0;      // sum of n^3 over the Gram blocks
};

// One two-site TDVP step (DmrgEngine::tdvp_step).
struct TdvpStats {
  double energy = 0;  // <psi|H|psi> / <psi|psi> of the last two-site tensor of the step
  double max_truncation_error = 0;
  int64_t max_bond_dim = 0;
  double seconds = 0;
  double log_norm = 0;  // ln of the norm factor the step applied (before renormalisation)
  int matvecs = 0;
  // host wall time of the phases (stream-synchronised): plan construction,
  // Krylov exponentials (with the two-site contraction and block copies),
  // splits, environment updates
  double t_plan = 0, t_exp = 0, t_svd = 0, t_env = 0;
};

struct SweepStats {
  double energy = 0;
  double max_truncation_error = 0;
  int64_t max_bond_dim = 0;
  double seconds = 0;
  int matvecs = 0, matvecs_fp64 = 0, steps = 0, unconverged = 0;
  // per-phase wall time (host clock around stream-synchronised phases)
  double t_plan = 0, t_eig = 0, t_svd = 0, t_env = 0;
};

class DmrgEngine {
 public:
  DmrgEngine(Context* ctx, DmrgConfig cfg);
  int n_sites() const { return n_; }
  Sym sym() const { return cfg_.sym; }
  const QIndex& phys(int i) const { return cfg_.mpo[i].phys; }
  // Initial MPS: bonds[0..N], site data in block-table order (host).
  void set_state(const std::vector<QIndex>& bonds, const std::vector<std::vector<double>>& site_data);
  // Right-canonicalise (centre at site 0, norm 1) and build the right environments.
  void prepare();
  SweepStats sweep();  // one left-to-right + right-to-left sweep
  // run_with_cancel_flag (sweep/mod.rs:274-291): while *flag reads non-zero,
  // sweep() and sweep_single_site() throw CancelledError before their next
  // step (null clears the flag).  The state stays a valid MPS; prepare()
  // again before the next sweep.
  void set_cancel_flag(const int32_t* flag) { cancel_ = flag; }
  // Subspace expansion of the single-site sweep (the expansion of
  // TdvpStabilizationConfig, ARCHITECTURE.md:1735-1790: expansion_alpha,
  // expansion_vectors).  With alpha > 0 every centre move that a one-site
  // solve follows perturbs the two-site tensor theta = A_i B_{i+1} by
  // alpha |theta| r_perp / |r_perp|, r_perp = H theta with its theta component
  // removed, and splits it with the two-site rule at
  // min(schedule D, current D + max_new) (max_new <= 0: the schedule's D):
  // directions the Hamiltonian couples to enter the bond, which a plain
  // single-site sweep can never grow.  alpha = 0 (the default) keeps the
  // untruncated, bond-preserving centre move.  The perturbation leaves the
  // bases of the bonds it passed O(alpha) away from the ground state's, so the
  // energy settles ~alpha^2 |E| above it; with `adaptive` (the reference's
  // adaptive_expansion) alpha drops tenfold after every sweep whose energy
  // change is within 10 alpha^2 max(1, |E|), and expansion stops below 1e-10.
  void set_subspace_expansion(double alpha, int64_t max_new, bool adaptive);
  double subspace_expansion_alpha() const { return expand_alpha_; }
  double energy() const { return energy_; }
  int sweeps_done() const { return sweeps_; }
  const std::vector<QIndex>& bonds() const { return bonds_; }
  const BlockTable& site_table(int i) const { return site_[i].t; }
  std::vector<double> site_data(int i) const;
  const double* site_device(int i) const { return site_[i].d.get(); }
  Context* context() const { return ctx_; }
  const std::vector<StepStats>& last_steps() const { return steps_; }
  // One second-order two-site TDVP step of psi <- exp(-tau H) psi (real tau:
  // imaginary-time evolution), the state renormalised after every local
  // update.  Left-to-right then right-to-left over every bond: the two-site
  // tensor evolves by exp(-tau/2 H2), is split (truncated to max_bond, the
  // sweep's cutoff), and, away from the ends, the new centre site evolves back
  // by exp(+tau/2 H1).  Krylov exponentials of krylov_dim vectors (at most,
  // with adaptive_tol > 0: the early stop of exp_krylov_generic).  Needs
  // prepare() (centre at site 0) and leaves the centre at site 0.
  TdvpStats tdvp_step(double tau, int64_t max_bond, int krylov_dim, double krylov_tol, double adaptive_tol = 0.0);
  // One left-to-right + right-to-left single-site sweep: at every site the
  // lowest eigenpair of H1 = L W R (the sweep's eigensolver, warm-started
  // from the site tensor) replaces the site, and the centre moves on by an
  // untruncated split of the site with its neighbour (the bond dimension
  // cannot grow).  Needs prepare() (centre at site 0); leaves it at site 0.
  SweepStats sweep_single_site();
  // psi <- O_0 psi for an operator O on site 0 of definite charge dq
  // (O[so][si] != 0 only where q(so) = q(si) + dq; rows = out, columns = in,
  // both in the sorted order of the site's physical QIndex).  Every bond
  // from 1 on shifts its charges by dq (dimensions unchanged), the state is
  // renormalised with the centre at site 0 and the environments rebuilt;
  // returns |O psi|^2 for the normalised psi.  Site 0 carries no
  // Jordan-Wigner string, so fermionic O act as given.
  double apply_site0_operator(const std::vector<double>& op, QN dq);

 private:
  struct Site {
    BlockTable t;
    DBuf<double> d;
  };
  struct Env {
    BlockTable t;
    DBuf<double> d;
  };
  StepStats step(int i, bool lr);
  // The two-site H_eff plan of sites (i, i+1): the one cached for i when the
  // environments' block layouts are those it was built for (rebound to the
  // current environment buffers), else a new one that replaces it.
  HeffPlan& two_site_plan(int i);
  // theta = A[i] A[i+1] into x (psi layout of `psi`)
  void contract_two_site(int i, const BlockTable& psi, double* x);
  // SVD of the two-site vector, truncation and write-back of A[i], A[i+1].
  void split_two_site(int i, bool lr, const BlockTable& psi, const double* x, int64_t max_bond, double cutoff,
                      double lambda_floor, StepStats* st);
  // site i <- exp(alpha H1) site i / |...| with H1 = L[i] W_i R[i+1]; returns
  // ln |exp(alpha H1) site i| - ln |site i| and adds matvecs and phase times to *ts
  double evolve_one_site(int i, double alpha, int krylov_dim, double tol, double adaptive_tol, TdvpStats* ts);
  // H1 = L[i] W_i R[i+1] as a two-site plan with a trivial second site, and
  // the block copies between site i and that plan's vector
  struct OneSite;
  OneSite one_site_plan(int i, Prec prec);
  StepStats step_single_site(int i, bool lr);
  // Move the orthogonality centre to site 0 by untruncated splits (the
  // two-site sweep leaves it at site 1) and build R[1].
  void centre_to_site0();
  void grow_left(int i);   // L[i+1] from L[i], A[i]
  void grow_right(int i);  // R[i] from R[i+1], A[i]
  Env boundary_env(const QIndex& bond, const QIndex& w);
  QIndex trivial_bond(QN q) const;

  Context* ctx_;
  DmrgConfig cfg_;
  int n_ = 0;
  std::vector<QIndex> bonds_;
  std::vector<Site> site_;
  std::vector<Env> L_, R_;
  struct CachedPlan {
    BlockTable l, r;  // layouts of the environments the plan was built for
    std::unique_ptr<HeffPlan> plan;
  };
  std::vector<CachedPlan> plan_cache_;
  double energy_ = 0;
  int sweeps_ = 0;
  int centre_ = -1;  // site of the orthogonality centre (-1 before prepare)
  std::vector<StepStats> steps_;
  SweepStats* cur_ = nullptr;
  const int32_t* cancel_ = nullptr;
  double expand_alpha_ = 0.0;
  int64_t expand_max_new_ = 0;
  bool expand_adaptive_ = true;
  double expand_e_prev_ = 0.0;
  bool expand_have_prev_ = false;
  void check_cancel();
};

// <bra|ket> of two MPS over the same physical legs: the transfer matrix
// E[q](bra, ket), block diagonal in the bond charge, carried from the left
// end to the right by two grouped contractions per site,
//   T = E[q] A_ket[s](q, q'),  E'[q'] = sum_{s, q} A_bra[s](q, q')^T T.
// The state norm is not divided out.  Reference: mps_overlap, named in the
// Green's function plan of tk-dmft/src/spectral/tdvp.rs:85-99.
double mps_overlap(const DmrgEngine& bra, const DmrgEngine& ket);

}  // namespace tk
