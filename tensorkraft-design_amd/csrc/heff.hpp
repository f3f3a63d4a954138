// Two-site effective Hamiltonian H_eff*psi = L * W_i * W_j * psi * R on block-sparse
// tensors, planned once per bond shape and applied many times inside the
// Krylov solver.
//
// Reference counterpart: build_heff_two_site + its matvec closure
// (tk-dmrg/src/environments/mod.rs:491-602; spec'd form with a backend
// argument in techspec/6_tech-spec_tk-dmrg.md:543-588).  The reference fills a
// dense dim x dim matrix; here the same sum is evaluated sector by sector:
//   stage 1  T1[w_l, s_i, s_j](a'_L, a_R) = L_{w_l}(a'_L, a_L) psi[s_i, s_j](a_L, a_R)
//   stage 2  the MPO pair sum  WW = sum_{w_m} W_i[s_i, s'_i, w_l, w_m] W_j[s_j, s'_j, w_m, w_r]
//            applied as coefficients on T1 blocks (fused into stage 3's A operand)
//   stage 3  y[s'_i, s'_j](a'_L, a'_R) = sum_{w_r} (sum WW * T1)(a'_L, a_R) R_{w_r}(a'_R, a_R)^T
// Layouts (all row-major, FlatBlockStorage order of ascending packed key):
//   psi/y   (s_i, s_j, a_L, a_R), dirs [Out, Out, Out, In], flux 0
//   L, R    (bra, w, ket),        dirs [In, In, Out],        flux 0  i.e. q(bra) = q(ket) - q(w)
//   W       dense (s_in, s_out, w_l, w_r); nonzeros obey q(s_in)+q(w_l) = q(s_out)+q(w_r)
#pragma once
#include <memory>
#include <vector>

#include "contract.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "sym.hpp"
#include "tune.hpp"

namespace tk {

enum class Prec : int32_t { FP64 = 0, FP32 = 1, BF16X3 = 2 };

struct MpoSite {
  QIndex phys;  // physical basis (the same on s_in and s_out)
  QIndex wl, wr;
  std::vector<double> w;  // dense phys.total^2 * wl.total * wr.total
  double at(int64_t si, int64_t so, int64_t a, int64_t b) const {
    return w[((si * phys.total + so) * wl.total + a) * wr.total + b];
  }
};

// Sparse W_i W_j product keyed by target (s'_i, s'_j, w_r) in full-index terms.
struct PairSource {
  int32_t si, sj, wl;
  double coef;
};
struct PairMpo {
  int64_t di, dj, wr_total;
  std::vector<int64_t> begin;  // per target index (s'_i*dj + s'_j)*wr_total + w_r
  std::vector<PairSource> src;
  double nnz_flops = 0;
};
PairMpo make_pair_mpo(const MpoSite& wi, const MpoSite& wj);
// Check that every nonzero of W respects the charge rule; throws ArgError.
void check_mpo_symmetry(const MpoSite& w);

// Full-index -> (sector, position) for a QIndex.
struct IndexMap {
  std::vector<int32_t> sector, within;
  explicit IndexMap(const QIndex& q);
};

struct HeffStats {
  double flops = 0;      // algorithmic flops of one apply
  double bytes = 0;      // algorithmic HBM bytes of one apply (at the compute precision)
  double s1_flops = 0, s1_bytes = 0, s3_flops = 0, s3_bytes = 0;  // per launch
  double s1_bytes_fp64 = 0, s3_bytes_fp64 = 0;                     // the FP64 refinement launches
  int64_t n_psi_blocks = 0, dim = 0, t1_elems = 0;
  int64_t stage1_items = 0, stage3_items = 0, stage1_tasks = 0, stage3_terms = 0;
};

class HeffPlan {
 public:
  // L and R are device pointers in FlatBlockStorage layout of `env_l`/`env_r`.
  HeffPlan(Context* ctx, Prec prec, const BlockTable& env_l, const double* L, const BlockTable& env_r,
           const double* R, const MpoSite& wi, const MpoSite& wj);
  const BlockTable& psi() const { return psi_; }
  const HeffStats& stats() const { return stats_; }
  Prec prec() const { return prec_; }
  // y = H_eff x; x, y device FP64 vectors of length psi().total (y overwritten).
  void apply(const double* x, double* y);
  // The same product replayed from a HIP graph of the apply's launches,
  // built at the first call (after any tuning) with explicit kernel nodes;
  // later calls re-point the nodes' x and y arguments and launch the graph.
  // Split-BF16 plans run apply().
  void apply_graph(const double* x, double* y);
  // y = H_eff x with every contraction in FP64 regardless of the plan precision.
  void apply_fp64(const double* x, double* y);
  ~HeffPlan();
  HeffPlan(const HeffPlan&) = delete;
  HeffPlan& operator=(const HeffPlan&) = delete;
  // The two launches of apply(), separately (for per-stage timing).
  void stage1(const double* x);
  void stage3(double* y);
  // d = diag(H_eff), FP64.
  void diagonal(double* d);
  // Re-bind environment data (same tables) without re-planning.
  void rebind(const double* L, const double* R);
  // The tiling each stage runs with (tune.hpp): the cached or measured one
  // when the context tunes, else the default.
  TileChoice tiling(int stage) const { return stage == 1 ? tiling1_ : tiling3_; }

 private:
  void build(const BlockTable& env_l, const BlockTable& env_r, const MpoSite& wi, const MpoSite& wj);
  Context* ctx_;
  Prec prec_;
  BlockTable psi_;
  const double* L_;
  const double* R_;
  int64_t l_total_, r_total_;
  ContractList h1_, h3_;
  DeviceContractList d1_, d3_;
  DBuf<double> t1_;
  DBuf<float> t1f_, xf_, Lf_, Rf_;
  bool low_ready_ = false;
  // diagonal terms: per psi slice, a list of (l_off, l_step, r_off, r_step, coef)
  struct DiagTerm { int64_t l_off, r_off; int32_t l_step, r_step; double coef; };
  struct DiagSlice { int64_t y_off; int32_t m, n, t_begin, t_end; };
  std::vector<DiagSlice> dslices_;
  std::vector<DiagTerm> dterms_;
  DBuf<char> ddev_;
  HeffStats stats_;
  uint64_t sig1_ = 0, sig3_ = 0;
  TileChoice tiling1_, tiling3_;
  bool tune_pending_ = false;
  // graph replay of apply(): nodes (with their argument slots), the index
  // of the argument that is x or y per node (-1: neither), the instance
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<KernelNode> gnodes_;
  std::vector<hipGraphNode_t> ghandles_;
  std::vector<int> gx_arg_, gy_arg_;
  // FP64 graphs: the slots holding L (stage 1's A) and R (stage 3's B), -1
  // elsewhere; a rebind re-points them instead of rebuilding the graph
  std::vector<int> gl_arg_, gr_arg_;
  // set the slot args[i] of every graph node i (-1: none) to `to`
  void patch_graph(const std::vector<int>& args, const void* to);
  void build_graph(const double* x, double* y);
  void drop_graph();
  void ensure_low();
  void set_tiling(const TileChoice& c1, const TileChoice& c3, int64_t t1_total);
  void launch_stage(int stage, const DeviceContractList& d, const double* x, double* y);
  void tune(const double* x, double* y);
};

void launch_heff_diag(Context* ctx, const void* slices, int n_slices, const void* terms, const double* L,
                      const double* R, double* d);

}  // namespace tk
