/*
 * tk_hip.h — C ABI of the MI355X block-sparse DMRG hot path.
 *
 * This is the drop-in boundary for tensorkraft's linear-algebra backend traits
 * (tk-linalg/src/traits.rs:25-205) and the H_eff producer of tk-dmrg
 * (tk-dmrg/src/environments/mod.rs:491-602).  A Rust `DeviceHip` would bind
 * these symbols from build.rs; the C++ host classes in
 * tensorkraft-design_amd/csrc/backend.hpp implement the mirrored traits on top
 * of them.
 *
 * Conventions
 *   - Every function returns a tk_status (0 = TK_OK); the message of the last
 *     failure on a context is returned by tk_last_error(ctx) (or
 *     tk_last_error(NULL) for calls that take no context).
 *   - Data pointers named *_dev are device pointers owned by the caller; all
 *     device work is enqueued on the context's stream and is asynchronous
 *     unless stated otherwise.  tk_ctx_sync waits for it.
 *   - One context per host thread.  Plans are owned by their creator and must
 *     be destroyed before their context.
 *   - Block-sparse data use FlatBlockStorage order
 *     (tk-symmetry/src/flat_storage.rs:95-140): blocks sorted by ascending
 *     packed sector key (tk-symmetry/src/sector_key.rs:31-55), each block
 *     row-major, blocks back to back.
 *   - Quantum numbers are passed as pairs (a, b): U1 uses (q, 0), U1U1 uses
 *     (N_up, N_dn).
 */
#ifndef TK_HIP_H
#define TK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t tk_status;
enum {
  TK_OK = 0,
  TK_ERR_INVALID_ARG = 1,        /* LinAlgError::DimensionMismatch and argument errors (tk-linalg/src/error.rs:16-21) */
  TK_ERR_SVD_CONVERGENCE = 2,    /* LinAlgError::SvdConvergence / SvdConvergenceError (tk-linalg/src/error.rs:6-9, results.rs:47) */
  TK_ERR_HIP = 3,                /* device runtime failure (the CudaError{op,detail} slot, tk-linalg/src/error.rs:29-32) */
  TK_ERR_NOT_CONVERGED = 4,      /* eigensolver did not reach its tolerance */
  TK_ERR_CANCELLED = 5,          /* DmrgError::Cancelled (tk-dmrg/src/error.rs:53) */
  TK_ERR_SYMMETRY = 6,           /* SymmetryError: flux rule, sector not found, QN range */
  TK_ERR_CAPACITY = 7,           /* caller-provided output array too small */
  TK_ERR_INTERNAL = 8
};
enum { TK_PREC_FP64 = 0, TK_PREC_FP32 = 1, TK_PREC_BF16X3 = 2 };
enum { TK_SYM_U1 = 1, TK_SYM_U1U1 = 2 };
enum { TK_DIR_IN = 0, TK_DIR_OUT = 1 }; /* LegDirection (tk-symmetry/src/quantum_number.rs) */

typedef struct tk_ctx tk_ctx;
typedef struct tk_heff_plan tk_heff_plan;

/* QIndex (tk-symmetry/src/sector_key.rs:95-158): sectors need not be sorted;
 * they are sorted by quantum number on use, as QIndex::new does. */
typedef struct {
  int32_t n_sectors;
  const int32_t* qn;   /* n_sectors x 2 */
  const int64_t* dims; /* n_sectors, all > 0 */
} tk_qindex;

/* MPO site tensor W[s_in][s_out][w_l][w_r], dense over the full index ranges
 * (tk-dmrg/src/mpo/mod.rs:36-93, leg order of mpo/mod.rs:3).  Nonzeros must
 * obey q(s_in) + q(w_l) = q(s_out) + q(w_r). */
typedef struct {
  tk_qindex phys;
  tk_qindex wl;
  tk_qindex wr;
  const double* w; /* host, phys^2 * wl * wr values */
} tk_mpo_site;

/* Block-sparse environment (bra, w, ket) with every flux-allowed block stored:
 * q(bra) = q(ket) - q(w).  Replaces the dense Environment of
 * tk-dmrg/src/environments/mod.rs:17-37 (layout env:90-95 per block). */
typedef struct {
  tk_qindex bond; /* bra and ket basis */
  tk_qindex w;    /* MPO bond basis */
  const double* data_dev;
} tk_env;

/* ---- version, context, plumbing ---------------------------------------- */
const char* tk_version(void);
/* One HIP stream + device memory pool.  Replaces the backend value
 * (DeviceFaer, tk-linalg/src/device/faer.rs) that every trait call carries. */
tk_status tk_ctx_create(int32_t device, tk_ctx** out);
tk_status tk_ctx_destroy(tk_ctx* ctx);
tk_status tk_ctx_sync(tk_ctx* ctx);
const char* tk_last_error(tk_ctx* ctx);
tk_status tk_device_alloc(tk_ctx* ctx, int64_t bytes, void** out_dev);
tk_status tk_device_free(tk_ctx* ctx, void* p_dev);
/* Synchronous copies (return after the data has arrived). */
tk_status tk_memcpy_h2d(tk_ctx* ctx, void* dst_dev, const void* src_host, int64_t bytes);
tk_status tk_memcpy_d2h(tk_ctx* ctx, void* dst_host, const void* src_dev, int64_t bytes);

/* Kernel timing (instrumentation; no counterpart in the reference).  When
 * enabled, the H_eff stage kernels and the eigensolver kernels of the SVD are
 * bracketed by HIP events on the context's stream; tk_ctx_timing_read waits
 * for them and returns, per kernel name, the launches, the summed event time
 * and the summed algorithmic bytes and flops, then clears the record.
 * Enabling or disabling clears it too. */
typedef struct {
  char name[40];
  int64_t launches;
  double ms;
  double bytes;
  double flops;
} tk_kernel_timing;
tk_status tk_ctx_timing(tk_ctx* ctx, int32_t enable);
tk_status tk_ctx_timing_read(tk_ctx* ctx, int32_t max, tk_kernel_timing* out, int32_t* n_out);

/* Contraction tiling autotuner (no counterpart in the reference, which
 * batches sector GEMMs by dimension without a measured choice,
 * ARCHITECTURE.md:1196-1218).  Each grouped contraction of an H_eff plan can
 * run as 32 x 32 wavefront tiles, 32 x 128 workgroup tiles sharing the A
 * chunk through LDS, or 64 x 64 workgroup tiles with both chunks in LDS (for
 * outputs from 32, 48 or 64 rows and columns on).  With mode 1 a plan whose
 * stage signature (precision, operand layout, log2 buckets of the block
 * sizes, terms and sources per output, half-octaves of the flops) has no cached choice times every
 * candidate with HIP events at its first apply and caches the fastest for
 * the process; mode 0 (the default) runs the 32 x 128 form.  Results of the
 * forms differ only in rounding. */
typedef struct {
  int32_t device;
  int32_t wg_kind;   /* 0 wavefront tiles, 1 32 x 128, 2 64 x 64 */
  int32_t tile_min;  /* 64 x 64 form: smallest output rows and columns it takes */
  int32_t pad;
  uint64_t signature;
  double us_best, us_default;  /* measured launch time of the choice and of the 32 x 128 form */
} tk_tune_entry;
tk_status tk_ctx_tune(tk_ctx* ctx, int32_t mode);
tk_status tk_tune_table(int32_t max, tk_tune_entry* out, int32_t* n_out);
tk_status tk_tune_save(const char* path);
tk_status tk_tune_load(const char* path);
tk_status tk_tune_clear(void);

/* ---- sector tables (host only) ------------------------------------------ */
/* QIndex::new / iter_sectors / total_dim (tk-symmetry/src/sector_key.rs:108-158):
 * sectors sorted by quantum number, prefix offsets.  Outputs have n_sectors
 * entries (qn_out: 2 per sector). */
tk_status tk_qindex_layout(int32_t sym, const tk_qindex* q, int32_t* qn_out, int64_t* dims_out,
                           int64_t* offsets_out, int64_t* total_out);
/* PackedSectorKey::pack (tk-symmetry/src/sector_key.rs:31-46) over `rank`
 * quantum numbers: leg i at bits [i*W, (i+1)*W), W = 8 (U1) or 16 (U1U1). */
tk_status tk_pack_key(int32_t sym, int32_t rank, const int32_t* qn, uint64_t* key_out);
/* enumerate_valid_sectors (tk-symmetry/src/flux.rs:43-68) + BlockSparseTensor::zeros
 * key sort (tk-symmetry/src/block_sparse.rs:40-84).  Writes up to `capacity`
 * blocks: keys, element offsets, per-block sector positions (rank entries,
 * indices into each leg's sorted QIndex) and the total element count. */
tk_status tk_block_table(int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                         const int32_t* flux, int64_t capacity, int64_t* n_blocks, uint64_t* keys,
                         int64_t* offsets, int32_t* sectors, int64_t* total);
/* fuse_legs offset map (tk-symmetry/src/block_sparse.rs:444-611): the fused
 * QIndex (sorted) and, per sub-sector combination in leg-by-leg order with the
 * later leg fastest, the fused sector position and the offset inside it. */
tk_status tk_fuse_map(int32_t sym, int32_t n_legs, const tk_qindex* legs, const int32_t* dirs,
                      int64_t capacity, int32_t* n_fused, int32_t* fused_qn, int64_t* fused_dims,
                      int64_t* n_combos, int32_t* combo_fused_sector, int64_t* combo_offset);

/* ---- block-sparse data movement (FP64, device data) ---------------------- */
/* Data are in the tk_block_table layout of the given legs (every flux-allowed
 * block, keys ascending, row-major blocks).  Synchronous. */
/* BlockSparseTensor::permute (tk-symmetry/src/block_sparse.rs:400-431):
 * leg k of the result is leg perm[k] of the input (QIndex and direction);
 * keys repacked and re-sorted; out_dev in the layout of the permuted legs. */
tk_status tk_block_permute(tk_ctx* ctx, int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                           const int32_t* flux, const int32_t* perm, const double* in_dev, int64_t in_elems,
                           double* out_dev, int64_t out_elems);
/* BlockSparseTensor::fuse_legs(start..end) (block_sparse.rs:444-611): out_dev
 * in the layout of legs [0,start), the fused leg (tk_fuse_map's QIndex of legs
 * [start,end), direction Incoming), legs [end,rank). */
tk_status tk_block_fuse(tk_ctx* ctx, int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                        const int32_t* flux, int32_t start, int32_t end, const double* in_dev, int64_t in_elems,
                        double* out_dev, int64_t out_elems);
/* BlockSparseTensor::split_leg (block_sparse.rs:622-793): the inverse of
 * tk_block_fuse for the same leg description; in_dev in the fused layout,
 * out_dev in the layout of the original legs. */
tk_status tk_block_split(tk_ctx* ctx, int32_t sym, int32_t rank, const tk_qindex* legs, const int32_t* dirs,
                         const int32_t* flux, int32_t start, int32_t end, const double* in_dev, int64_t in_elems,
                         double* out_dev, int64_t out_elems);

/* ---- LinAlgBackend / SparseLinAlgBackend (FP64, device data) ------------ */
/* LinAlgBackend::gemm (tk-linalg/src/traits.rs:114-121; faer.rs:591-639):
 * C = alpha A B + beta C with MatRef strides (tk-core/src/matview.rs:12-33);
 * each operand needs a unit stride on one axis, else TK_ERR_INVALID_ARG
 * (NonContiguousForBlas, tk-linalg/src/error.rs:24-27).  Synchronous. */
tk_status tk_gemm(tk_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* a_dev, int64_t rs_a,
                  int64_t cs_a, const double* b_dev, int64_t rs_b, int64_t cs_b, double beta, double* c_dev,
                  int64_t rs_c, int64_t cs_c);

/* No counterpart in the reference: tk_gemm for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops).  Same
 * arguments, results bit-identical to tk_gemm's; the call's work list (16
 * bytes per 32 x 32 output tile) lives in a static 4 MiB device array per
 * GPU instead of a per-call allocation, so outputs above about 262,000
 * tiles are rejected with TK_ERR_INVALID_ARG.  Calls that overlap on one GPU
 * (contexts of the same device on different threads) race on that array.
 * Synchronous. */
tk_status tk_gemm_noalloc(tk_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* a_dev,
                          int64_t rs_a, int64_t cs_a, const double* b_dev, int64_t rs_b, int64_t cs_b, double beta,
                          double* c_dev, int64_t rs_c, int64_t cs_c);

/* Rank-2 block-sparse matrix, legs [Incoming, Outgoing] (block_gemm_realistic.rs:26). */
typedef struct {
  tk_qindex rows;
  tk_qindex cols;
  int32_t flux[2];
  const double* data_dev; /* FlatBlockStorage order */
} tk_block_matrix;

/* SparseLinAlgBackend::block_gemm (traits.rs:190-205; faer.rs:390-575):
 * the output's block table (blocks receiving at least one sector task, keys
 * ascending, flux = a.flux + b.flux) ... */
tk_status tk_block_gemm_table(int32_t sym, const tk_block_matrix* a, const tk_block_matrix* b, int64_t capacity,
                              int64_t* n_blocks, uint64_t* keys, int64_t* offsets, int64_t* total);
/* ... and the product into c_dev (that table's layout).  Equal output keys
 * accumulate in the fixed order (flops descending, key_a, key_b).  Synchronous. */
tk_status tk_block_gemm(tk_ctx* ctx, int32_t sym, const tk_block_matrix* a, const tk_block_matrix* b,
                        double* c_dev);
/* SparseLinAlgBackend::spmv (faer.rs:336-388): y = A x, y zeroed first; x and
 * y over the full sorted index ranges of cols / rows.  Synchronous. */
tk_status tk_spmv(tk_ctx* ctx, int32_t sym, const tk_block_matrix* a, const double* x_dev, double* y_dev);
/* LinAlgBackend::svd_truncated (traits.rs:38-88; faer.rs:127-240): row-major
 * m x n A, min(m, n) <= 2048 (else TK_ERR_INVALID_ARG); keeps at most
 * max_rank singular values with sigma >= cutoff * sigma_max (the
 * truncate_svd! rule, faer.rs:135-147).  Two Gram levels resolve sigma to
 * ~eps sigma_max^2 / sigma above 3.2e-5 sigma_max and to ~1e-13 sigma_max
 * below, so the default cutoff 1e-12 holds except for values within a factor
 * ~3 of it; values no level resolves count as zero.  U (m x rank), Vt
 * (rank x n) row-major on the device (capacity m x min(m, n, max_rank) and
 * min(m, n, max_rank) x n), s (host) descending, truncation_error = sum of
 * discarded sigma^2 (results.rs:18-23). */
tk_status tk_svd_truncated(tk_ctx* ctx, int64_t m, int64_t n, const double* a_dev, int64_t max_rank, double cutoff,
                           double* u_dev, double* s_host, double* vt_dev, int64_t* rank, double* truncation_error);
/* truncate_svd (tk-dmrg/src/truncation/mod.rs:56-127): the full SVD of a
 * row-major rows x cols matrix (min <= 2048) cut by TruncationConfig
 * (truncation/mod.rs:12-30): keep while count < max_bond_dim and sigma >=
 * svd_cutoff * sigma_max, then at least min(min_bond_dim, min(rows, cols)).
 * U (rows x bond_dim_new) and Vt (bond_dim_new x cols) on the device, s
 * (host, bond_dim_new values), n_discarded = min(rows, cols) - bond_dim_new,
 * truncation_error relative: 1 - sum kept sigma^2 / sum sigma^2. */
tk_status tk_truncate_svd(tk_ctx* ctx, int64_t rows, int64_t cols, const double* a_dev, int64_t max_bond_dim,
                          double svd_cutoff, int64_t min_bond_dim, double* u_dev, double* s_host, double* vt_dev,
                          int64_t* bond_dim_new, int64_t* n_discarded, double* truncation_error);
/* The selection rule of truncate_svd (truncation/mod.rs:72-98) over the
 * singular values of several blocks at once, as the two-site split applies
 * it (host only, no device work).  sigma holds n_blocks runs of counts[b]
 * values, each descending.  Candidates are ordered by sigma descending (ties:
 * block, then position); keep while count < max_bond and sigma >= cutoff *
 * sigma_max and sigma^2 >= sigma2_floor * sigma_max^2; when degeneracy_tol > 0
 * the cut moves up to the start of a group sigma_k - sigma_{k+1} <=
 * degeneracy_tol * sigma_max (unless nothing would be kept); then at least
 * min(min_bond, candidates).  Outputs: keep[b] per block, the totals and the
 * relative truncation error.  The reference's single-matrix rule is
 * n_blocks = 1, degeneracy_tol = 0, sigma2_floor = 0. */
tk_status tk_truncation_select(int32_t n_blocks, const int64_t* counts, const double* sigma, int64_t max_bond,
                               double cutoff, int64_t min_bond, double degeneracy_tol, double sigma2_floor,
                               int64_t* keep, int64_t* n_kept, int64_t* n_discarded, double* truncation_error);
/* The block-diagonal truncated SVD of the two-site split (the reference's
 * truncate_svd on the whole matricised two-site tensor, sweep/mod.rs:401-421,
 * here per middle charge): blocks b = 0..n_blocks-1 of rows[b] x cols[b]
 * (rows[b] <= 2048) lie back to back, row-major, in x_dev; the selection of
 * tk_truncation_select over all blocks' singular values.  Outputs: keep[b];
 * sigma_host: the kept values block after block (each run descending); u_dev:
 * per block the rows[b] x keep[b] left singular vectors, back to back; c_dev:
 * per block the keep[b] x cols[b] matrix U_b^T X_b (= diag(sigma) Vt, the
 * next centre of a left-to-right step), back to back.  The capacities bound
 * sigma_host (values), u_dev and c_dev (elements); TK_ERR_CAPACITY when one
 * is too small.  Synchronous. */
tk_status tk_block_svd_truncate(tk_ctx* ctx, int32_t n_blocks, const int64_t* rows, const int64_t* cols,
                                const double* x_dev, int64_t max_bond, double cutoff, int64_t min_bond,
                                double degeneracy_tol, double sigma2_floor, int64_t* keep, double* sigma_host,
                                int64_t sigma_capacity, double* u_dev, int64_t u_capacity, double* c_dev,
                                int64_t c_capacity, int64_t* n_kept, int64_t* n_discarded, double* truncation_error);
/* LinAlgBackend::eigh_lowest (traits.rs:134-138), all pairs: eigenvalues
 * ascending (host), eigenvectors as columns of a row-major n x n (device);
 * n <= 2048. */
tk_status tk_eigh(tk_ctx* ctx, int64_t n, const double* a_dev, double* evals_host, double* evecs_dev);
/* LinAlgBackend::eigh_lowest (traits.rs:134-138; faer.rs:243-290): the k
 * lowest eigenpairs of a symmetric n x n (n <= 2048), eigenvalues ascending
 * (host, k), eigenvectors as the columns of a row-major n x k (device).
 * k > n is LinAlgError::EighKTooLarge: TK_ERR_INVALID_ARG. */
tk_status tk_eigh_lowest(tk_ctx* ctx, int64_t n, const double* a_dev, int64_t k, double* evals_host,
                         double* evecs_dev);
/* LinAlgBackend::qr (traits.rs:148; faer.rs:292-324): thin QR of a row-major
 * m x n A: Q (m x k, orthonormal columns) and R (k x n, upper triangular),
 * k = min(m, n), row-major on the device (QrResult, results.rs:34-40);
 * Householder reflectors, so R's diagonal may be negative.  A is not
 * modified.  Synchronous. */
tk_status tk_qr(tk_ctx* ctx, int64_t m, int64_t n, const double* a_dev, double* q_dev, double* r_dev);
/* LinAlgBackend::regularized_svd_inverse (traits.rs:150-183): out (n x m,
 * device) = V diag(s_i / (s_i^2 + delta^2)) U^T for U (m x k) and Vt (k x n)
 * on the device and s (k, host); delta > 0. */
tk_status tk_regularized_svd_inverse(tk_ctx* ctx, int64_t m, int64_t n, int64_t k, const double* s_host,
                                     const double* u_dev, const double* vt_dev, double delta, double* out_dev);
/* No counterpart in the reference: q[i] = a[i] / b[i] for device arrays of n
 * doubles, formed as the eigensolver's Sturm counts and inverse iteration
 * form their quotients (hardware reciprocal of b and one Newton step, then
 * the product with a), for callers that check that quotient's error against
 * IEEE division.  Synchronous. */
tk_status tk_fast_quotient(tk_ctx* ctx, const double* a_dev, const double* b_dev, double* q_dev, int64_t n);
/* BondDimensionSchedule::warmup (tk-dmrg/src/truncation/mod.rs:137-151):
 * round(d_init * r^i) (half away from zero), r = (d_max / d_init)^(1/n),
 * clamped to [1, d_max], i = 0..n (n + 1 values; n = 0 gives [d_max]).
 * bond_dim_at_sweep (truncation/mod.rs:163-171) is the value at the sweep
 * index, the last one beyond the end: tk_dmrg_config.bond_schedule takes the
 * list with that meaning.  Host only. */
tk_status tk_bond_schedule_warmup(int64_t d_init, int64_t d_max, int64_t n_warmup_sweeps, int64_t capacity,
                                  int64_t* dims, int64_t* n_dims);

/* IterativeEigensolver::lowest_eigenpair (tk-dmrg/src/eigensolver/mod.rs:38-55)
 * with DavidsonSolver semantics (eigensolver/mod.rs:300-516) on a caller
 * matvec over device vectors: matvec(x_dev, y_dev, user) must compute y = A x
 * on the context's stream (or synchronise before returning).  x_dev holds the
 * start vector and receives the normalised eigenvector; diag_dev (nullable)
 * is the preconditioner diagonal. */
typedef void (*tk_matvec_fn)(const double* x_dev, double* y_dev, void* user);
tk_status tk_davidson(tk_ctx* ctx, int64_t n, tk_matvec_fn matvec, void* user, const double* diag_dev, double tol,
                      int32_t max_subspace, int32_t max_iter, double* x_dev, double* eigenvalue, double* residual,
                      int32_t* matvecs);
/* LanczosSolver::lowest_eigenpair (tk-dmrg/src/eigensolver/mod.rs:57-297) over
 * the same matvec callback: full reorthogonalisation, at most max_krylov
 * (<= 512) Krylov vectors per pass, restart from the Ritz vector up to
 * max_restarts times; stops when the Ritz residual estimate is below tol.
 * x_dev: start vector in, normalised eigenvector out; *residual is the
 * explicit |A x - eigenvalue x|; *matvecs counts every product (including the
 * final residual). */
tk_status tk_lanczos(tk_ctx* ctx, int64_t n, tk_matvec_fn matvec, void* user, double tol, int32_t max_krylov,
                     int32_t max_restarts, double* x_dev, double* eigenvalue, double* residual, int32_t* matvecs);
/* exp_krylov_f64 + mat_exp_times_vec (tk-dmrg/src/tdvp/mod.rs:197-382):
 * out_dev = exp(alpha A) v_dev for real alpha over the same matvec callback.
 * min(krylov_dim, n) Arnoldi vectors (krylov_dim <= 48), happy breakdown at
 * h_{j+1,j} < tol, the small exponential by the reference's Taylor series with
 * scaling and squaring; |v| < 1e-30 gives zero.  out_dev may alias v_dev.
 * adaptive_tol > 0 (not in the reference; 0 keeps its fixed dimension) stops
 * after step m once h_{m+1,m} |c_m| <= adaptive_tol |c| for c = exp(alpha H_m)
 * |v| e_1.  *krylov_used (nullable) receives the Arnoldi vectors used. */
tk_status tk_exp_krylov(tk_ctx* ctx, int64_t n, tk_matvec_fn matvec, void* user, const double* v_dev, double alpha,
                        int32_t krylov_dim, double tol, double adaptive_tol, double* out_dev, int32_t* krylov_used);

/* ---- two-site effective Hamiltonian ------------------------------------- */
/* build_heff_two_site (tk-dmrg/src/environments/mod.rs:491-602;
 * techspec/6_tech-spec_tk-dmrg.md:561-572).  psi layout (s_i, s_j, a_L, a_R)
 * with dirs [Out, Out, Out, In], flux 0.  The plan keeps pointers to the
 * environment data; they must stay valid while the plan is used. */
tk_status tk_heff_plan_create(tk_ctx* ctx, int32_t prec, int32_t sym, const tk_env* left, const tk_env* right,
                              const tk_mpo_site* wi, const tk_mpo_site* wj, tk_heff_plan** out);
tk_status tk_heff_plan_destroy(tk_heff_plan* plan);
/* The psi block table the plan uses (keys ascending, offsets). */
tk_status tk_heff_psi_table(tk_heff_plan* plan, int64_t capacity, int64_t* n_blocks, uint64_t* keys,
                            int64_t* offsets, int64_t* total);
/* y = H_eff x (the matvec closure, environments/mod.rs:592-599). */
tk_status tk_heff_apply(tk_heff_plan* plan, const double* x_dev, double* y_dev);
/* The same product from a HIP graph of the apply's kernel launches (built at
 * the first call; later calls re-point its x and y and replay it).  The
 * Davidson and Lanczos solvers of the sweep multiply through this path, in
 * every precision. */
tk_status tk_heff_apply_graph(tk_heff_plan* plan, const double* x_dev, double* y_dev);
/* The tiling stage 1 or 3 of the plan runs with (after its first apply when
 * the context tunes): wg_kind and tile_min as in tk_tune_entry. */
tk_status tk_heff_tiling(tk_heff_plan* plan, int32_t stage, int32_t* wg_kind, int32_t* tile_min);
/* y = H_eff x with FP64 contractions whatever the plan precision (refinement). */
tk_status tk_heff_apply_fp64(tk_heff_plan* plan, const double* x_dev, double* y_dev);
/* d = diag(H_eff), the Davidson preconditioner (eigensolver/mod.rs:425-435). */
tk_status tk_heff_diag(tk_heff_plan* plan, double* d_dev);
/* Algorithmic flops and HBM bytes of one apply; counts[0..5] = psi blocks,
 * dim, T1 elements, stage-1 items, stage-3 items, stage-3 terms. */
tk_status tk_heff_stats(tk_heff_plan* plan, double* flops, double* bytes, int64_t* counts);
/* Per-launch algorithmic work: out[0..3] = stage-1 flops, stage-1 bytes,
 * stage-3 flops, stage-3 bytes. */
tk_status tk_heff_stage_stats(tk_heff_plan* plan, double* out);
/* Times `reps` applies with HIP events on the plan's stream; returns the mean
 * per apply in microseconds and the mean per launch of each stage. */
tk_status tk_heff_time(tk_heff_plan* plan, const double* x_dev, double* y_dev, int32_t reps, double* us_apply,
                       double* us_stage1, double* us_stage3);

/* ---- two-site DMRG ------------------------------------------------------ */
typedef struct tk_dmrg tk_dmrg;

/* DMRGConfig + TruncationConfig + DavidsonSolver (tk-dmrg/src/sweep/mod.rs:124-162,
 * truncation/mod.rs:12-30, eigensolver/mod.rs:300-323). */
typedef struct {
  int32_t prec;                 /* TK_PREC_*: precision of the H_eff contractions */
  int32_t sym;                  /* TK_SYM_* */
  int32_t n_sites;
  int32_t n_schedule;
  const int64_t* bond_schedule; /* max bond dimension per sweep (BondDimensionSchedule, truncation/mod.rs:131-172) */
  double svd_cutoff;            /* keep sigma >= cutoff * sigma_max */
  int64_t min_bond_dim;
  double degeneracy_tol;        /* no cut inside |s_k - s_{k+1}| <= tol * s_max */
  double davidson_tol;          /* absolute residual 2-norm */
  int32_t max_subspace;
  int32_t max_iter;
  int32_t precondition;         /* diagonal preconditioner on/off */
  double switch_rel;            /* low-precision -> FP64 refinement switch */
  int32_t eigensolver;          /* TK_EIG_DAVIDSON (DMRGConfig's default) or TK_EIG_LANCZOS */
  int32_t max_krylov;           /* Lanczos vectors per pass (<= 512; 0: 100) */
  double low_tail;              /* with prec below FP64, Davidson corrections are multiplied in
                                   that precision once the residual is below low_tail *
                                   davidson_tol (0: 1000; negative: never) */
} tk_dmrg_config;
enum { TK_EIG_DAVIDSON = 0, TK_EIG_LANCZOS = 1 };

typedef struct {
  double energy;                /* last step's eigenvalue (sweep/mod.rs:301) */
  double max_truncation_error;
  double seconds;               /* wall time of the sweep, stream-synchronised */
  double t_plan, t_eig, t_svd, t_env;
  int64_t max_bond_dim;
  int32_t matvecs, matvecs_fp64, steps, unconverged;
} tk_sweep_stats;

typedef struct {
  int32_t site, left_to_right, iterations, matvecs, matvecs_fp64, converged;
  double energy, truncation_error, residual;
  int64_t bond_dim;
  int64_t dim;       /* two-site vector length */
  int64_t svd_nmax;  /* largest Gram block of the SVD */
  double svd_n3;     /* sum of n^3 over the Gram blocks */
} tk_step_stats;

/* DMRGEngine::new (sweep/mod.rs:234-265) with the MPO given site by site
 * (mpo[i].wr must equal mpo[i+1].wl; the ends have dimension 1). */
tk_status tk_dmrg_create(tk_ctx* ctx, const tk_dmrg_config* cfg, const tk_mpo_site* mpo, tk_dmrg** out);
tk_status tk_dmrg_destroy(tk_dmrg* d);
/* Initial MPS: n_sites+1 bond bases and per-site data (host, block-table order
 * of (s, a_L, a_R), dirs [Out, Out, In]).  Bond 0 must be {0:1}, bond N the
 * total charge with dimension 1. */
tk_status tk_dmrg_set_state(tk_dmrg* d, const tk_qindex* bonds, const double* const* site_data);
/* Right-canonicalise and build the right environments (Environments::build_from_scratch,
 * environments/mod.rs:117-155, on a canonical MPS). */
tk_status tk_dmrg_prepare(tk_dmrg* d);
/* One left-to-right + right-to-left sweep (SweepSchedule::standard, sweep/mod.rs:59-81). */
tk_status tk_dmrg_sweep(tk_dmrg* d, tk_sweep_stats* out);
/* The cancel flag of DMRGEngine::run_with_cancel_flag (sweep/mod.rs:274-291):
 * while the int32 at `flag` (host memory, may be null to clear) reads
 * non-zero, tk_dmrg_sweep and tk_dmrg_sweep_single_site stop before their
 * next two-site (or one-site) step and return TK_ERR_CANCELLED
 * (DmrgError::Cancelled, tk-dmrg/src/error.rs:53).  Another thread may set
 * the flag at any time; it is read with a relaxed atomic load.  Steps
 * already done are kept (the state stays a valid MPS); the engine must be
 * prepared again (tk_dmrg_prepare) before the next sweep. */
tk_status tk_dmrg_set_cancel_flag(tk_dmrg* d, const int32_t* flag);
/* Subspace expansion for tk_dmrg_sweep_single_site (the expansion_alpha /
 * expansion_vectors of TdvpStabilizationConfig, ARCHITECTURE.md:1735-1790,
 * applied to the one-site sweep, whose reference step sweep/mod.rs:451-509
 * never grows a bond).  With alpha > 0 every centre move that a one-site
 * solve follows adds alpha |theta| r / |r| to the two-site tensor
 * theta = A_i B_{i+1}, r = H_eff theta with its theta component removed, and
 * splits the result with the two-site truncation rule at
 * min(schedule D, current D + max_new) (max_new <= 0: the schedule's D), so
 * the bond takes up the directions H couples to.  The perturbation leaves the
 * bases O(alpha) away from the ground state's (the energy settles about
 * alpha^2 |E| above it); with adaptive != 0 (adaptive_expansion) alpha drops
 * tenfold after every single-site sweep whose energy change is within
 * 10 alpha^2 max(1, |E|), and expansion stops below 1e-10.  alpha = 0 (the
 * default) restores the untruncated centre move; alpha < 0 or non-finite is
 * TK_ERR_INVALID_ARG. */
tk_status tk_dmrg_set_subspace_expansion(tk_dmrg* d, double alpha, int64_t max_new, int32_t adaptive);
/* The expansion's current alpha (after the adaptive decreases). */
tk_status tk_dmrg_subspace_expansion_alpha(tk_dmrg* d, double* alpha);
/* dmrg_step_single_site (tk-dmrg/src/sweep/mod.rs:451-509) over a whole sweep.
 * The reference solves the one-site problem but neither writes the eigenvector
 * back nor moves the centre; here every site is replaced by the lowest
 * eigenpair of L W R (the config's eigensolver, warm-started from the site)
 * and the centre moves on by an untruncated split with the neighbour, so the
 * bond dimensions never grow.  Left to right then right to left; needs
 * tk_dmrg_prepare.  Statistics as for tk_dmrg_sweep (truncation error 0 up to
 * numerical zeros). */
tk_status tk_dmrg_sweep_single_site(tk_dmrg* d, tk_sweep_stats* out);
/* Per-step statistics of the last sweep (StepResult, sweep/mod.rs:35-50). */
tk_status tk_dmrg_steps(tk_dmrg* d, int32_t capacity, int32_t* n, tk_step_stats* out);
/* Current bond basis (sorted) and site tensor data. */
tk_status tk_dmrg_bond(tk_dmrg* d, int32_t bond, int32_t capacity, int32_t* n_sectors, int32_t* qn, int64_t* dims);
tk_status tk_dmrg_site(tk_dmrg* d, int32_t site, int64_t capacity, int64_t* n, double* data);

/* Step 2 of compute_greens_function_tdvp (tk-dmft/src/spectral/tdvp.rs:55-108, where
 * it is described but not implemented): psi <- O psi for an operator O on site 0
 * of definite charge dq[0..1] (U1: dq[1] = 0).  op is d x d row-major (row = out,
 * column = in) in the ascending charge order of site 0's physical QIndex.  Bonds
 * 1..N shift their charges by dq, the state is renormalised with the centre at
 * site 0 and the environments are rebuilt; *norm2 = |O psi|^2 of the normalised
 * input.  Site 0 carries no Jordan-Wigner string, so c and c+ act as given. */
tk_status tk_dmrg_apply_site_op(tk_dmrg* d, const double* op, const int32_t* dq, double* norm2);

/* mps_overlap (named in the Green's function plan of
 * tk-dmft/src/spectral/tdvp.rs:85-99): *out = <bra|ket> of the two engines'
 * current states (same site count, symmetry and physical legs; norms not
 * divided out).  Both engines must live on one device. */
tk_status tk_dmrg_overlap(tk_dmrg* bra, tk_dmrg* ket, double* out);

/* TdvpStepResult (tk-dmrg/src/tdvp/mod.rs:20-33) plus the step's energy and norm factor. */
typedef struct {
  double energy;                /* <psi|H|psi>/<psi|psi> of the step's last two-site tensor */
  double max_truncation_error;
  double seconds;
  double log_norm;              /* ln of the norm factor exp(-tau H) applied (before renormalisation) */
  double t_plan, t_exp, t_svd, t_env;  /* phase wall times: plans, Krylov exponentials, splits, environments */
  int64_t max_bond_dim;
  int32_t matvecs, pad;
} tk_tdvp_stats;
/* TdvpDriver::step (tk-dmrg/src/tdvp/mod.rs:91-141), which in the reference
 * leaves the MPS unchanged: here one second-order two-site TDVP step of
 * psi <- exp(-tau H) psi for real tau (imaginary time), renormalised.  Bonds
 * left to right then right to left: exp(-tau/2 H2) on the two-site tensor
 * (exp_krylov_f64, tdvp/mod.rs:197-272), split truncated to max_bond (<= 0:
 * the last bond-schedule value) with the config's cutoff, and away from the
 * ends exp(+tau/2 H1) on the new centre site.  krylov_adaptive_tol as in
 * tk_exp_krylov (0: the reference's fixed Krylov dimension).  Needs
 * tk_dmrg_prepare; the centre is at site 0 afterwards. */
tk_status tk_dmrg_tdvp_step(tk_dmrg* d, double tau, int64_t max_bond, int32_t krylov_dim, double krylov_tol,
                            double krylov_adaptive_tol, tk_tdvp_stats* out);

/* ---- DMFT impurity pieces (tk-dmft) ---------------------------------------
 * BathParameters::hybridization_function (tk-dmft/src/impurity/bath.rs:61-83),
 * batched over problems: out[(p * n_z + j) * 2 + {0,1}] = Re, Im Delta_p(z_j)
 * with z_j = z_re[j] + i z_im[j] (the reference's omega + i*broadening, or
 * i*omega_n on the Matsubara axis).  params (n_problems x 2 n_bath):
 * eps[0..n_bath), v[0..n_bath).  geometry TK_BATH_STAR: sum_k v_k^2/(z - eps_k);
 * TK_BATH_CHAIN: v_0^2/(z - eps_0 - v_1^2/(z - eps_1 - ...)), the impurity
 * coupled to a tight-binding chain (v_k: hopping of chain sites k-1, k).
 * Device pointers; synchronous. */
#define TK_BATH_STAR 0
#define TK_BATH_CHAIN 1
tk_status tk_hybridization(tk_ctx* ctx, int32_t geometry, int32_t n_problems, int32_t n_bath,
                           const double* params_dev, int32_t n_z, const double* z_re_dev, const double* z_im_dev,
                           double* out_dev);
/* No reference counterpart (the DMFT loop's least-squares bath update, used
 * when no learned proposal is taken): per problem, star parameters minimising
 * sum_j |Delta(i wn_j) - delta_j|^2 by Levenberg-Marquardt.  params_dev holds
 * the start on entry and the fit on exit; cost_host (n_problems) the final
 * cost.  n_z <= 256, n_bath <= 8.  Device pointers except cost_host. */
tk_status tk_bath_fit(tk_ctx* ctx, int32_t n_problems, int32_t n_bath, int32_t n_z, const double* wn_dev,
                      const double* delta_dev, double* params_dev, int32_t max_iter, double* cost_host);
/* lanczos_tridiagonalize (tk-dmft/src/impurity/discretize.rs:54-274), the
 * recursion on weighted points (x_j, w_j >= 0): eps[k] = alpha_k, v[0] =
 * sqrt(sum w), v[k] = beta_k; the reference's grid form passes x = omega,
 * w = -Im Delta(omega) d_omega / pi.  Host arrays. */
tk_status tk_lanczos_tridiagonalize(int32_t n, const double* x, const double* w, int32_t n_bath, double* eps,
                                    double* v);
/* BroydenState (tk-dmft/src/loop/mixing.rs:40-150): update(x, f, alpha). */
typedef struct tk_broyden tk_broyden;
tk_status tk_broyden_create(int32_t history_depth, tk_broyden** out);
tk_status tk_broyden_update(tk_broyden* b, int32_t dim, const double* x, const double* f, double alpha, double* out,
                            int32_t* n_updates);
tk_status tk_broyden_destroy(tk_broyden* b);
/* Impurity Green's function as a continued fraction (no reference
 * counterpart; the reference's G(t) route is TDVP): Lanczos with full
 * reorthogonalisation on the n x n CSR matrix (rowptr n+1, col, val; device)
 * from v0 (device).  beta[0] = |v0|, alpha[j], beta[j+1] (host, m each);
 * *steps = number of alphas. */
tk_status tk_krylov_tridiag(tk_ctx* ctx, int64_t n, const int64_t* rowptr_dev, const int32_t* col_dev,
                            const double* val_dev, const double* v0_dev, int32_t m, double* alpha, double* beta,
                            int32_t* steps);
/* No reference counterpart (the learned bath proposal): FP32 multilayer
 * perceptron dims[0] -> ... -> dims[n_layers], tanh hidden layers, linear
 * output, Glorot-uniform weights from `seed`.  Parameters are [W_1 (row-major
 * d_0 x d_1), b_1, W_2, b_2, ...]. */
typedef struct tk_mlp tk_mlp;
tk_status tk_mlp_create(tk_ctx* ctx, int32_t n_layers, const int32_t* dims, uint64_t seed, tk_mlp** out);
tk_status tk_mlp_destroy(tk_mlp* m);
tk_status tk_mlp_n_params(tk_mlp* m, int64_t* n);
tk_status tk_mlp_get_params(tk_mlp* m, float* host);
tk_status tk_mlp_set_params(tk_mlp* m, const float* host);
/* y (batch x dims[last]) = net(x (batch x dims[0])); device pointers. */
tk_status tk_mlp_forward(tk_mlp* m, int32_t batch, const float* x_dev, float* y_dev);
/* Mean squared error against t (device) and its gradient wrt every
 * parameter (grad_host may be null); the gradient is kept for tk_mlp_adam. */
tk_status tk_mlp_loss_grad(tk_mlp* m, int32_t batch, const float* x_dev, const float* t_dev, float* loss,
                           float* grad_host);
tk_status tk_mlp_adam(tk_mlp* m, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif /* TK_HIP_H */
