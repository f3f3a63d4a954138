// FP64 CPU oracle for the block-sparse two-site DMRG hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain restatement of tensorkraft's
// algorithms for checking the HIP path; it is called only by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product library.  It shares no code with tensorkraft-design_amd/.
//
// What each part restates, and what pins it (parity status per function):
//   sector tables ......... tk-symmetry/src/{builtins,sector_key,flux,block_sparse}.rs
//                           pinned: flux.rs:259-329 counts 7/141/51, sector_key.rs:246-254
//                           offsets, builtins.rs round trips, brute-force enumeration
//   fuse map .............. tk-symmetry/src/block_sparse.rs:444-611
//                           pinned: fuse->split round trips (block_sparse.rs:1030-1115)
//   block_gemm ............ tk-linalg/src/device/faer.rs:390-575
//                           pinned: block_gemm_realistic.rs LCG goldens vs dense GEMM
//   H_eff apply ........... tk-dmrg/src/environments/mod.rs:491-602 (block-sparse form)
//                           pinned: the reference's dense loop restated verbatim at tiny D,
//                           Hermiticity, 2-site Heisenberg singlet -3/4
//   env update ............ tk-dmrg/src/environments/mod.rs:225-483 (block-sparse form)
//   SVD / truncation ...... tk-dmrg/src/truncation/mod.rs:56-127, faer.rs:127-181
//                           pinned: truncation/mod.rs:185-236 (sigma 3,2; error 1/14), LAPACK via numpy
//   Davidson .............. tk-dmrg/src/eigensolver/mod.rs:325-601
//                           pinned: eigensolver/mod.rs:660-801 known answers
//   DMRG sweep ............ tk-dmrg/src/sweep/mod.rs:274-448 (with the defect fixes of the design)
//                           pinned: exact diagonalisation, free-fermion closed form
#pragma once
#include <cstdint>

extern "C" {

typedef struct {
  int32_t n_sectors;
  const int32_t* qn;
  const int64_t* dims;
} or_qindex;

// All flux-allowed sector tuples by brute force over every tuple, sorted by key.
int or_block_table(int32_t sym, int32_t rank, const or_qindex* legs, const int32_t* dirs, const int32_t* flux,
                   int64_t capacity, int64_t* n_blocks, uint64_t* keys, int64_t* offsets, int64_t* total);
int or_pack_key(int32_t sym, int32_t rank, const int32_t* qn, uint64_t* key);
int or_fuse_map(int32_t sym, int32_t n_legs, const or_qindex* legs, const int32_t* dirs, int64_t capacity,
                int32_t* n_fused, int32_t* fused_qn, int64_t* fused_dims, int64_t* n_combos,
                int32_t* combo_fused_sector, int64_t* combo_offset);

// Rank-2 block_gemm: C = A * B with A, B given by (row, col) qindices + flux,
// data in key order.  Output table written as keys/offsets/data.
int or_block_gemm(int32_t sym, const or_qindex* a_rows, const or_qindex* a_cols, const int32_t* a_flux,
                  const double* a_data, const or_qindex* b_rows, const or_qindex* b_cols, const int32_t* b_flux,
                  const double* b_data, int64_t capacity, int64_t* n_blocks, uint64_t* keys, int64_t* offsets,
                  double* c_data, int64_t* c_total);

// y = H_eff x, block-sparse FP64.  Environments (bond, w, bond) with every
// allowed block (q(bra) = q(ket) - q(w)), data in key order; W dense over
// the caller's index order.  nthreads <= 0 uses all cores.
int or_heff_apply(int32_t sym, const or_qindex* phys_i, const or_qindex* phys_j, const or_qindex* bond_l,
                  const or_qindex* bond_r, const or_qindex* w_l, const or_qindex* w_m, const or_qindex* w_r,
                  const double* wi, const double* wj, const double* L, const double* R, const double* x, double* y,
                  int32_t nthreads);
// diag(H_eff)
int or_heff_diag(int32_t sym, const or_qindex* phys_i, const or_qindex* phys_j, const or_qindex* bond_l,
                 const or_qindex* bond_r, const or_qindex* w_l, const or_qindex* w_m, const or_qindex* w_r,
                 const double* wi, const double* wj, const double* L, const double* R, double* d);
// The reference's dense H_eff (environments/mod.rs:505-599) restated loop by
// loop on densified environments, then y = H x in the dense (s_i, s_j, a_L, a_R)
// layout (full-index order).  Tiny sizes only: dim^2 doubles.
// truncate_svd's selection rule (truncation/mod.rs:56-98) over the singular
// values of n_blocks blocks (counts[b] values each, descending), with the
// design's degeneracy rule (deg_tol > 0) and numerical-zero floor on sigma^2.
int or_truncation_select(int32_t n_blocks, const int64_t* counts, const double* sigma, int64_t max_bond,
                         double cutoff, int64_t min_bond, double deg_tol, double floor, int64_t* keep,
                         int64_t* n_kept, int64_t* n_disc, double* terr);
int or_heff_dense_reference(int32_t sym, const or_qindex* phys_i, const or_qindex* phys_j, const or_qindex* bond_l,
                            const or_qindex* bond_r, const or_qindex* w_l, const or_qindex* w_m,
                            const or_qindex* w_r, const double* wi, const double* wj, const double* L,
                            const double* R, const double* x_dense, double* y_dense);
}
