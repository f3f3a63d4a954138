// Grouped block contraction: the work list format shared by the host planner
// and the persistent MFMA kernel.
//
// One launch computes a set of output blocks
//     C_o = sum_{t in terms(o)} ( sum_{s in srcs(t)} coef_s * op(A_s) ) * op(B_t)
// where every operand is a row-major slice (base offset + leading dimension)
// of one of three device buffers.  This covers, in one format,
//   * the rank-2 block_gemm of SparseLinAlgBackend (tk-linalg/src/device/faer.rs:390-575),
//     with one term and one source per sector GEMM task;
//   * the two grouped-GEMM stages of H_eff*psi with the MPO W_i W_j application
//     fused into the A-operand of the second stage (the A-side source sum);
//   * the environment updates (tk-dmrg/src/environments/mod.rs:225-483).
// Accumulation order is fixed by the term/source order, so results are
// deterministic; each output tile is owned by exactly one wavefront.
#pragma once
#include <cstdint>
#include <vector>

namespace tk {

struct GOut {
  int64_t c_off;
  int32_t m, n, ldc;
  int32_t t_begin, t_end;
  int32_t pad;
};
struct GTerm {
  int64_t b_off;
  int32_t ldb, k;
  int32_t s_begin, s_end;
};
struct GSrc {
  int64_t a_off;
  int32_t lda, pad;
  double coef;
};
// One wavefront work item: a 32x32 tile of output block `out`.
struct GItem {
  int32_t out, m0, n0, cost;
};

// Operand transposition flags of one launch.
enum ContractFlags : int { kTransA = 1, kTransB = 2, kTransC = 4 };

// Narrowest output (columns) tiled into workgroup items when a list is wide.
constexpr int kWideMinN = 96;

// Workgroup item kinds: 32 x 128 items with the A chunk shared through LDS
// (kWgWide), or 64 x 64 items with both operand chunks staged in LDS
// (kWgTile64).
enum WgKind : int { kWgNone = 0, kWgWide = 1, kWgTile64 = 2 };

// Host-side list builder.
struct ContractList {
  std::vector<GOut> outs;
  std::vector<GTerm> terms;
  std::vector<GSrc> srcs;
  std::vector<GItem> items;
  // With wg_kind != kWgNone, the outputs the kind takes (kWgWide: at least
  // kWideMinN columns; kWgTile64: at least tile_min rows and columns) are
  // tiled into workgroup items (`witems`); items[0, n_narrow) then tile only
  // the other outputs, and the rest of `items` tiles the workgroup outputs in
  // 32 x 32 for kernels without that workgroup form.
  int wg_kind = kWgNone;
  int tile_min = 48;
  std::vector<GItem> witems;
  int n_narrow = 0;
  int flags = 0;
  double beta = 0;   // C = sum + beta * C (0: C is not read)
  double flops = 0;  // algorithmic flops of the list (2*m*n*k per term + 2*m*k per extra source)

  int begin_out(int64_t c_off, int m, int n, int ldc) {
    outs.push_back({c_off, m, n, ldc, int32_t(terms.size()), int32_t(terms.size()), 0});
    return int(outs.size()) - 1;
  }
  void begin_term(int64_t b_off, int ldb, int k) {
    terms.push_back({b_off, ldb, k, int32_t(srcs.size()), int32_t(srcs.size())});
    outs.back().t_end = int32_t(terms.size());
  }
  void add_src(int64_t a_off, int lda, double coef) {
    srcs.push_back({a_off, lda, 0, coef});
    terms.back().s_end = int32_t(srcs.size());
  }
  // Drop a term that ended up with no sources.
  void drop_empty_term() {
    if (!terms.empty() && terms.back().s_begin == terms.back().s_end) {
      terms.pop_back();
      outs.back().t_end = int32_t(terms.size());
    }
  }
  // Tile every output into 32x32 wavefront items, longest first (LPT order,
  // tk-linalg/src/tasks.rs:172-174) with a stable tie-break on position.
  void finalize();
  // Checks every operand slice against the element counts of the three
  // buffers (throws std::out_of_range); run before a list is first launched.
  void validate(int64_t a_elems, int64_t b_elems, int64_t c_elems) const;
  void clear() {
    outs.clear(); terms.clear(); srcs.clear(); items.clear(); witems.clear();
    n_narrow = 0;
    flops = 0;
  }
};

}  // namespace tk
