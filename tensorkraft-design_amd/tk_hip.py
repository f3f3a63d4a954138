"""ctypes binding of libtk_hip.so (the C ABI declared in include/tk_hip.h).

The library is the product: every call here runs the HIP path.  Loading fails
loudly when the shared object has not been built; there is no fallback.

Also holds the workload builders used by the benchmark and the tests:
Hubbard / Heisenberg MPO site tensors (basis and Jordan-Wigner conventions of
tk-dsl/src/operators.rs:31-200) and the synthetic U(1)xU(1) bond of the single
H_eff*psi microbenchmark.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtk_hip.so")

TK_OK = 0
TK_ERR_INVALID_ARG = 1
TK_ERR_CANCELLED = 5
TK_ERR_CAPACITY = 7
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
SYM_U1, SYM_U1U1 = 1, 2
DIR_IN, DIR_OUT = 0, 1
STATUS = {0: "TK_OK", 1: "TK_ERR_INVALID_ARG", 2: "TK_ERR_SVD_CONVERGENCE", 3: "TK_ERR_HIP",
          4: "TK_ERR_NOT_CONVERGED", 5: "TK_ERR_CANCELLED", 6: "TK_ERR_SYMMETRY", 7: "TK_ERR_CAPACITY",
          8: "TK_ERR_INTERNAL"}


class TkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class tk_qindex(C.Structure):
    _fields_ = [("n_sectors", C.c_int32), ("qn", C.POINTER(C.c_int32)), ("dims", C.POINTER(C.c_int64))]


class tk_mpo_site(C.Structure):
    _fields_ = [("phys", tk_qindex), ("wl", tk_qindex), ("wr", tk_qindex), ("w", C.POINTER(C.c_double))]


class tk_env(C.Structure):
    _fields_ = [("bond", tk_qindex), ("w", tk_qindex), ("data_dev", C.c_void_p)]


class tk_block_matrix(C.Structure):
    _fields_ = [("rows", tk_qindex), ("cols", tk_qindex), ("flux", C.c_int32 * 2), ("data_dev", C.c_void_p)]


MATVEC_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)


class tk_dmrg_config(C.Structure):
    _fields_ = [("prec", C.c_int32), ("sym", C.c_int32), ("n_sites", C.c_int32), ("n_schedule", C.c_int32),
                ("bond_schedule", C.POINTER(C.c_int64)), ("svd_cutoff", C.c_double), ("min_bond_dim", C.c_int64),
                ("degeneracy_tol", C.c_double), ("davidson_tol", C.c_double), ("max_subspace", C.c_int32),
                ("max_iter", C.c_int32), ("precondition", C.c_int32), ("switch_rel", C.c_double),
                ("eigensolver", C.c_int32), ("max_krylov", C.c_int32), ("low_tail", C.c_double)]


class tk_kernel_timing(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("launches", C.c_int64), ("ms", C.c_double), ("bytes", C.c_double),
                ("flops", C.c_double)]


class tk_tdvp_stats(C.Structure):
    _fields_ = [("energy", C.c_double), ("max_truncation_error", C.c_double), ("seconds", C.c_double),
                ("log_norm", C.c_double), ("t_plan", C.c_double), ("t_exp", C.c_double), ("t_svd", C.c_double),
                ("t_env", C.c_double), ("max_bond_dim", C.c_int64), ("matvecs", C.c_int32), ("pad", C.c_int32)]


class tk_sweep_stats(C.Structure):
    _fields_ = [("energy", C.c_double), ("max_truncation_error", C.c_double), ("seconds", C.c_double),
                ("t_plan", C.c_double), ("t_eig", C.c_double), ("t_svd", C.c_double), ("t_env", C.c_double),
                ("max_bond_dim", C.c_int64), ("matvecs", C.c_int32), ("matvecs_fp64", C.c_int32),
                ("steps", C.c_int32), ("unconverged", C.c_int32)]


class tk_tune_entry(C.Structure):
    _fields_ = [("device", C.c_int32), ("wg_kind", C.c_int32), ("tile_min", C.c_int32), ("pad", C.c_int32),
                ("signature", C.c_uint64), ("us_best", C.c_double), ("us_default", C.c_double)]


class tk_step_stats(C.Structure):
    _fields_ = [("site", C.c_int32), ("left_to_right", C.c_int32), ("iterations", C.c_int32),
                ("matvecs", C.c_int32), ("matvecs_fp64", C.c_int32), ("converged", C.c_int32),
                ("energy", C.c_double), ("truncation_error", C.c_double), ("residual", C.c_double),
                ("bond_dim", C.c_int64), ("dim", C.c_int64), ("svd_nmax", C.c_int64), ("svd_n3", C.c_double)]


_lib = None


def lib():
    """Load libtk_hip.so once; raise if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, U64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    pI32, pI64, pU64, pD = C.POINTER(I32), C.POINTER(I64), C.POINTER(U64), C.POINTER(D)
    sig = {
        "tk_version": (C.c_char_p, []),
        "tk_ctx_create": (I32, [I32, C.POINTER(P)]),
        "tk_ctx_destroy": (I32, [P]),
        "tk_ctx_sync": (I32, [P]),
        "tk_last_error": (C.c_char_p, [P]),
        "tk_device_alloc": (I32, [P, I64, C.POINTER(P)]),
        "tk_device_free": (I32, [P, P]),
        "tk_memcpy_h2d": (I32, [P, P, P, I64]),
        "tk_memcpy_d2h": (I32, [P, P, P, I64]),
        "tk_pack_key": (I32, [I32, I32, pI32, pU64]),
        "tk_qindex_layout": (I32, [I32, C.POINTER(tk_qindex), pI32, pI64, pI64, pI64]),
        "tk_block_table": (I32, [I32, I32, C.POINTER(tk_qindex), pI32, pI32, I64, pI64, pU64, pI64, pI32, pI64]),
        "tk_fuse_map": (I32, [I32, I32, C.POINTER(tk_qindex), pI32, I64, pI32, pI32, pI64, pI64, pI32, pI64]),
        "tk_heff_plan_create": (I32, [P, I32, I32, C.POINTER(tk_env), C.POINTER(tk_env), C.POINTER(tk_mpo_site),
                                      C.POINTER(tk_mpo_site), C.POINTER(P)]),
        "tk_heff_plan_destroy": (I32, [P]),
        "tk_heff_psi_table": (I32, [P, I64, pI64, pU64, pI64, pI64]),
        "tk_heff_apply": (I32, [P, P, P]),
        "tk_heff_apply_fp64": (I32, [P, P, P]),
        "tk_heff_diag": (I32, [P, P]),
        "tk_heff_stats": (I32, [P, pD, pD, pI64]),
        "tk_heff_stage_stats": (I32, [P, pD]),
        "tk_heff_time": (I32, [P, P, P, I32, pD, pD, pD]),
        "tk_gemm": (I32, [P, I64, I64, I64, D, P, I64, I64, P, I64, I64, D, P, I64, I64]),
        "tk_gemm_noalloc": (I32, [P, I64, I64, I64, D, P, I64, I64, P, I64, I64, D, P, I64, I64]),
        "tk_fast_quotient": (I32, [P, P, P, P, I64]),
        "tk_block_gemm_table": (I32, [I32, C.POINTER(tk_block_matrix), C.POINTER(tk_block_matrix), I64, pI64, pU64,
                                      pI64, pI64]),
        "tk_block_gemm": (I32, [P, I32, C.POINTER(tk_block_matrix), C.POINTER(tk_block_matrix), P]),
        "tk_spmv": (I32, [P, I32, C.POINTER(tk_block_matrix), P, P]),
        "tk_svd_truncated": (I32, [P, I64, I64, P, I64, D, P, pD, P, pI64, pD]),
        "tk_eigh": (I32, [P, I64, P, pD, P]),
        "tk_truncate_svd": (I32, [P, I64, I64, P, I64, D, I64, P, pD, P, pI64, pI64, pD]),
        "tk_truncation_select": (I32, [I32, pI64, pD, I64, D, I64, D, D, pI64, pI64, pI64, pD]),
        "tk_block_svd_truncate": (I32, [P, I32, pI64, pI64, P, I64, D, I64, D, D, pI64, pD, I64, P, I64, P, I64,
                                        pI64, pI64, pD]),
        "tk_eigh_lowest": (I32, [P, I64, P, I64, pD, P]),
        "tk_qr": (I32, [P, I64, I64, P, P, P]),
        "tk_regularized_svd_inverse": (I32, [P, I64, I64, I64, pD, P, P, D, P]),
        "tk_bond_schedule_warmup": (I32, [I64, I64, I64, I64, pI64, pI64]),
        "tk_dmrg_set_cancel_flag": (I32, [P, pI32]),
        "tk_dmrg_set_subspace_expansion": (I32, [P, D, I64, I32]),
        "tk_dmrg_subspace_expansion_alpha": (I32, [P, C.POINTER(C.c_double)]),
        "tk_davidson": (I32, [P, I64, MATVEC_FN, P, P, D, I32, I32, P, pD, pD, pI32]),
        "tk_lanczos": (I32, [P, I64, MATVEC_FN, P, D, I32, I32, P, pD, pD, pI32]),
        "tk_exp_krylov": (I32, [P, I64, MATVEC_FN, P, P, D, I32, D, D, P, pI32]),
        "tk_dmrg_tdvp_step": (I32, [P, D, I64, I32, D, D, C.POINTER(tk_tdvp_stats)]),
        "tk_dmrg_sweep_single_site": (I32, [P, C.POINTER(tk_sweep_stats)]),
        "tk_dmrg_apply_site_op": (I32, [P, pD, pI32, pD]),
        "tk_dmrg_overlap": (I32, [P, P, pD]),
        "tk_dmrg_create": (I32, [P, C.POINTER(tk_dmrg_config), C.POINTER(tk_mpo_site), C.POINTER(P)]),
        "tk_dmrg_destroy": (I32, [P]),
        "tk_dmrg_set_state": (I32, [P, C.POINTER(tk_qindex), C.POINTER(C.POINTER(D))]),
        "tk_dmrg_prepare": (I32, [P]),
        "tk_dmrg_sweep": (I32, [P, C.POINTER(tk_sweep_stats)]),
        "tk_dmrg_steps": (I32, [P, I32, pI32, C.POINTER(tk_step_stats)]),
        "tk_dmrg_bond": (I32, [P, I32, I32, pI32, pI32, pI64]),
        "tk_dmrg_site": (I32, [P, I32, I64, pI64, pD]),
        "tk_ctx_timing": (I32, [P, I32]),
        "tk_ctx_tune": (I32, [P, I32]),
        "tk_tune_table": (I32, [I32, C.POINTER(tk_tune_entry), pI32]),
        "tk_tune_save": (I32, [C.c_char_p]),
        "tk_tune_load": (I32, [C.c_char_p]),
        "tk_tune_clear": (I32, []),
        "tk_heff_tiling": (I32, [P, I32, pI32, pI32]),
        "tk_heff_apply_graph": (I32, [P, P, P]),
        "tk_block_permute": (I32, [P, I32, I32, C.POINTER(tk_qindex), pI32, pI32, pI32, P, I64, P, I64]),
        "tk_block_fuse": (I32, [P, I32, I32, C.POINTER(tk_qindex), pI32, pI32, I32, I32, P, I64, P, I64]),
        "tk_block_split": (I32, [P, I32, I32, C.POINTER(tk_qindex), pI32, pI32, I32, I32, P, I64, P, I64]),
        "tk_ctx_timing_read": (I32, [P, I32, C.POINTER(tk_kernel_timing), pI32]),
        "tk_hybridization": (I32, [P, I32, I32, I32, P, I32, P, P, P]),
        "tk_bath_fit": (I32, [P, I32, I32, I32, P, P, P, I32, pD]),
        "tk_lanczos_tridiagonalize": (I32, [I32, pD, pD, I32, pD, pD]),
        "tk_broyden_create": (I32, [I32, C.POINTER(P)]),
        "tk_broyden_update": (I32, [P, I32, pD, pD, D, pD, pI32]),
        "tk_broyden_destroy": (I32, [P]),
        "tk_krylov_tridiag": (I32, [P, I64, P, P, P, P, I32, pD, pD, pI32]),
        "tk_mlp_create": (I32, [P, I32, pI32, U64, C.POINTER(P)]),
        "tk_mlp_destroy": (I32, [P]),
        "tk_mlp_n_params": (I32, [P, pI64]),
        "tk_mlp_get_params": (I32, [P, C.POINTER(C.c_float)]),
        "tk_mlp_set_params": (I32, [P, C.POINTER(C.c_float)]),
        "tk_mlp_forward": (I32, [P, I32, P, P]),
        "tk_mlp_loss_grad": (I32, [P, I32, P, P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "tk_mlp_adam": (I32, [P, C.c_float, C.c_float, C.c_float, C.c_float]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = ["tk_version", "tk_ctx_create", "tk_ctx_destroy", "tk_ctx_sync", "tk_last_error", "tk_device_alloc",
            "tk_device_free", "tk_memcpy_h2d", "tk_memcpy_d2h", "tk_pack_key", "tk_qindex_layout", "tk_block_table", "tk_fuse_map",
            "tk_heff_plan_create", "tk_heff_plan_destroy", "tk_heff_psi_table", "tk_heff_apply",
            "tk_heff_apply_fp64", "tk_heff_diag", "tk_heff_stats", "tk_heff_stage_stats", "tk_heff_time",
            "tk_dmrg_create", "tk_dmrg_destroy", "tk_dmrg_set_state", "tk_dmrg_prepare", "tk_dmrg_sweep",
            "tk_dmrg_steps", "tk_dmrg_bond", "tk_dmrg_site", "tk_gemm", "tk_gemm_noalloc", "tk_fast_quotient", "tk_block_gemm_table",
            "tk_block_gemm",
            "tk_spmv", "tk_svd_truncated", "tk_eigh", "tk_truncate_svd", "tk_truncation_select",
            "tk_block_svd_truncate", "tk_eigh_lowest", "tk_qr", "tk_regularized_svd_inverse",
            "tk_bond_schedule_warmup", "tk_dmrg_set_cancel_flag", "tk_dmrg_set_subspace_expansion", "tk_dmrg_subspace_expansion_alpha", "tk_davidson", "tk_lanczos", "tk_exp_krylov", "tk_dmrg_tdvp_step",
            "tk_dmrg_sweep_single_site", "tk_dmrg_apply_site_op", "tk_dmrg_overlap",
            "tk_ctx_timing", "tk_ctx_tune", "tk_tune_table", "tk_tune_save", "tk_tune_load", "tk_tune_clear",
            "tk_heff_tiling", "tk_heff_apply_graph",
            "tk_ctx_timing_read", "tk_block_permute", "tk_block_fuse", "tk_block_split", "tk_hybridization",
            "tk_bath_fit", "tk_lanczos_tridiagonalize", "tk_broyden_create", "tk_broyden_update",
            "tk_broyden_destroy", "tk_krylov_tridiag", "tk_mlp_create", "tk_mlp_destroy", "tk_mlp_n_params",
            "tk_mlp_get_params", "tk_mlp_set_params", "tk_mlp_forward", "tk_mlp_loss_grad", "tk_mlp_adam"]


def tune_table():
    """The tuner's cache: [{device, wg_kind, tile_min, signature, us_best, us_default}]."""
    n = C.c_int32(0)
    lib().tk_tune_table(0, None, C.byref(n))
    buf = (tk_tune_entry * max(1, n.value))()
    _check(lib().tk_tune_table(len(buf), buf, C.byref(n)))
    return [{f: getattr(buf[i], f) for f, _ in tk_tune_entry._fields_ if f != "pad"} for i in range(n.value)]


def tune_save(path):
    _check(lib().tk_tune_save(os.fsencode(path)))


def tune_load(path):
    _check(lib().tk_tune_load(os.fsencode(path)))


def tune_clear():
    _check(lib().tk_tune_clear())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def truncation_select(sigmas, max_bond, cutoff, min_bond=0, degeneracy_tol=0.0, sigma2_floor=0.0):
    """The library's truncation rule (tk_truncation_select; truncate_svd's
    selection, truncation/mod.rs:72-98, over several blocks): (keep per
    block, n_kept, n_discarded, relative truncation error).  Host only."""
    counts = np.array([len(x) for x in sigmas], dtype=np.int64)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in sigmas])
                                if len(sigmas) else np.zeros(1))
    keep = np.zeros(max(1, len(sigmas)), dtype=np.int64)
    nk, nd, te = C.c_int64(0), C.c_int64(0), C.c_double(0)
    _check(lib().tk_truncation_select(len(sigmas), _ip(counts), _dp(flat), int(max_bond), float(cutoff),
                                      int(min_bond), float(degeneracy_tol), float(sigma2_floor), _ip(keep),
                                      C.byref(nk), C.byref(nd), C.byref(te)))
    return keep[:len(sigmas)].tolist(), nk.value, nd.value, te.value


class BondDimensionSchedule:
    """BondDimensionSchedule (tk-dmrg/src/truncation/mod.rs:131-172); warmup
    comes from the library (tk_bond_schedule_warmup)."""

    def __init__(self, dims):
        self.dims = [int(d) for d in dims]

    @classmethod
    def warmup(cls, d_init, d_max, n_warmup_sweeps):
        out = np.zeros(int(n_warmup_sweeps) + 1, dtype=np.int64)
        n = C.c_int64(0)
        _check(lib().tk_bond_schedule_warmup(int(d_init), int(d_max), int(n_warmup_sweeps), len(out), _ip(out),
                                             C.byref(n)))
        return cls(out[:n.value].tolist())

    @classmethod
    def fixed(cls, d):
        return cls([d])

    @classmethod
    def custom(cls, dims):
        return cls(dims)

    def bond_dim_at_sweep(self, sweep_index):
        if 0 <= sweep_index < len(self.dims):
            return self.dims[sweep_index]
        return self.dims[-1] if self.dims else 1


def _check(st, ctx=None):
    if st != TK_OK:
        msg = lib().tk_last_error(ctx).decode()
        raise TkError(st, msg)


# ---------------------------------------------------------------- qindex helpers
class QIndexArg:
    """Keeps the numpy arrays of a tk_qindex alive."""

    def __init__(self, sectors):
        self.sectors = [(tuple(q) if isinstance(q, (tuple, list)) else (int(q), 0), int(d)) for q, d in sectors]
        self.qn = np.ascontiguousarray(np.array([c for q, _ in self.sectors for c in q], dtype=np.int32))
        self.dims = np.ascontiguousarray(np.array([d for _, d in self.sectors], dtype=np.int64))
        self.c = tk_qindex(len(self.sectors), self.qn.ctypes.data_as(C.POINTER(C.c_int32)),
                           self.dims.ctypes.data_as(C.POINTER(C.c_int64)))

    @property
    def total(self):
        return int(self.dims.sum())


def sorted_sectors(sectors):
    """QIndex::new order (sector_key.rs:108-117): by quantum number."""
    return sorted(((tuple(q) if isinstance(q, (tuple, list)) else (int(q), 0)), int(d)) for q, d in sectors)


def qindex_layout(sym, sectors):
    """Product QIndex::new: sorted (qn, dim) list, offsets, total."""
    q = QIndexArg(sectors)
    n = len(q.sectors)
    qn = np.zeros(2 * n, dtype=np.int32)
    dims = np.zeros(n, dtype=np.int64)
    offs = np.zeros(n, dtype=np.int64)
    tot = C.c_int64(0)
    _check(lib().tk_qindex_layout(sym, C.byref(q.c), qn.ctypes.data_as(C.POINTER(C.c_int32)),
                                  dims.ctypes.data_as(C.POINTER(C.c_int64)), offs.ctypes.data_as(C.POINTER(C.c_int64)),
                                  C.byref(tot)))
    return [((int(qn[2 * i]), int(qn[2 * i + 1])), int(dims[i])) for i in range(n)], offs, tot.value


def block_table(sym, legs, dirs, flux=(0, 0)):
    """Product block table: (keys, offsets, sector positions, total)."""
    qs = [QIndexArg(l) for l in legs]
    arr = (tk_qindex * len(qs))(*[q.c for q in qs])
    d = np.array(dirs, dtype=np.int32)
    f = np.array(flux, dtype=np.int32)
    n = C.c_int64(0)
    tot = C.c_int64(0)
    L = lib()
    st = L.tk_block_table(sym, len(qs), arr, d.ctypes.data_as(C.POINTER(C.c_int32)),
                          f.ctypes.data_as(C.POINTER(C.c_int32)), 0, C.byref(n), None, None, None, C.byref(tot))
    if st not in (TK_OK, TK_ERR_CAPACITY):
        _check(st)
    cap = n.value
    keys = np.zeros(cap, dtype=np.uint64)
    offs = np.zeros(cap, dtype=np.int64)
    secs = np.zeros(cap * len(qs), dtype=np.int32)
    _check(L.tk_block_table(sym, len(qs), arr, d.ctypes.data_as(C.POINTER(C.c_int32)),
                            f.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n),
                            keys.ctypes.data_as(C.POINTER(C.c_uint64)), offs.ctypes.data_as(C.POINTER(C.c_int64)),
                            secs.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(tot)))
    return keys, offs, secs.reshape(cap, len(qs)), tot.value


def pack_key(sym, qns):
    a = np.array([c for q in qns for c in ((q, 0) if np.isscalar(q) else q)], dtype=np.int32)
    k = C.c_uint64(0)
    _check(lib().tk_pack_key(sym, len(qns), a.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(k)))
    return k.value


def fuse_map(sym, legs, dirs):
    qs = [QIndexArg(l) for l in legs]
    arr = (tk_qindex * len(qs))(*[q.c for q in qs])
    d = np.array(dirs, dtype=np.int32)
    nf = C.c_int32(0)
    nc = C.c_int64(0)
    cap = 1
    for l in legs:
        cap *= len(l)
    fq = np.zeros(2 * cap, dtype=np.int32)
    fd = np.zeros(cap, dtype=np.int64)
    cs = np.zeros(cap, dtype=np.int32)
    co = np.zeros(cap, dtype=np.int64)
    _check(lib().tk_fuse_map(sym, len(qs), arr, d.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(nf),
                             fq.ctypes.data_as(C.POINTER(C.c_int32)), fd.ctypes.data_as(C.POINTER(C.c_int64)),
                             C.byref(nc), cs.ctypes.data_as(C.POINTER(C.c_int32)),
                             co.ctypes.data_as(C.POINTER(C.c_int64))))
    nfv = nf.value
    fused = [((int(fq[2 * i]), int(fq[2 * i + 1])), int(fd[i])) for i in range(nfv)]
    return fused, cs[:nc.value].copy(), co[:nc.value].copy()


# ---------------------------------------------------------------- device context
class Context:
    """tk_ctx: one HIP stream + device pool on `device`."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        _check(self.L.tk_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def adopt(self, child):
        """Register an object that must be closed before the context."""
        self._children.add(child)

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.close()
            self.L.tk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timing(self, enable=True):
        """Enable (and clear) or disable the HIP-event timing of the H_eff and
        eigensolver kernels on this context's stream."""
        _check(self.L.tk_ctx_timing(self.h, 1 if enable else 0), self.h)

    def tune(self, on=True):
        """Measure the contraction tiling per signature at first use (tk_ctx_tune)."""
        _check(self.L.tk_ctx_tune(self.h, 1 if on else 0), self.h)

    def timing_read(self):
        """Per kernel: launches, event milliseconds, algorithmic bytes and flops
        recorded since timing was enabled (clears the record)."""
        buf = (tk_kernel_timing * 64)()
        n = C.c_int32(0)
        _check(self.L.tk_ctx_timing_read(self.h, len(buf), buf, C.byref(n)), self.h)
        return {buf[i].name.decode(): {"launches": buf[i].launches, "ms": buf[i].ms, "bytes": buf[i].bytes,
                                       "flops": buf[i].flops} for i in range(n.value)}

    def sync(self):
        _check(self.L.tk_ctx_sync(self.h), self.h)

    def alloc(self, nbytes):
        p = C.c_void_p()
        _check(self.L.tk_device_alloc(self.h, int(nbytes), C.byref(p)), self.h)
        return p.value

    def free(self, p):
        _check(self.L.tk_device_free(self.h, C.c_void_p(p)), self.h)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(max(arr.nbytes, 8))
        _check(self.L.tk_memcpy_h2d(self.h, C.c_void_p(p), arr.ctypes.data_as(C.c_void_p), arr.nbytes), self.h)
        return DeviceArray(self, p, arr.shape, arr.dtype)

    def empty(self, n, dtype=np.float64):
        dt = np.dtype(dtype)
        return DeviceArray(self, self.alloc(max(n * dt.itemsize, 8)), (n,), dt)


class DeviceArray:
    def __init__(self, ctx, ptr, shape, dtype):
        self.ctx, self.ptr, self.shape, self.dtype = ctx, ptr, tuple(shape), np.dtype(dtype)

    @property
    def nbytes(self):
        return int(np.prod(self.shape)) * self.dtype.itemsize

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx.L.tk_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                        self.nbytes), self.ctx.h)
        return out

    def free(self):
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = None


# ---------------------------------------------------------------- backend traits
class HipBackend:
    """LinAlgBackend / SparseLinAlgBackend on the device (tk-linalg/src/traits.rs:25-205).

    Methods take and return numpy arrays and move them through the context;
    the device work is the library's.
    """

    def __init__(self, ctx):
        self.ctx = ctx

    def gemm(self, alpha, a, b, beta, c):
        """C = alpha A B + beta C (row-major numpy operands; returns the new C)."""
        a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
        c = np.array(c, dtype=np.float64)
        m, k = a.shape
        k2, n = b.shape
        if k2 != k or c.shape != (m, n):
            raise ValueError("gemm: dimension mismatch")
        da, db, dc = self.ctx.to_device(a), self.ctx.to_device(b), self.ctx.to_device(c)
        _check(self.ctx.L.tk_gemm(self.ctx.h, m, n, k, alpha, C.c_void_p(da.ptr), k, 1, C.c_void_p(db.ptr), n, 1,
                                  beta, C.c_void_p(dc.ptr), n, 1), self.ctx.h)
        out = dc.get()
        for x in (da, db, dc):
            x.free()
        return out

    def gemm_strided(self, m, n, k, alpha, a_dev, rs_a, cs_a, b_dev, rs_b, cs_b, beta, c_dev, rs_c, cs_c):
        _check(self.ctx.L.tk_gemm(self.ctx.h, m, n, k, alpha, C.c_void_p(a_dev.ptr), rs_a, cs_a,
                                  C.c_void_p(b_dev.ptr), rs_b, cs_b, beta, C.c_void_p(c_dev.ptr), rs_c, cs_c),
               self.ctx.h)

    @staticmethod
    def _bm(sym, rows, cols, flux, dev, keep):
        r, c = QIndexArg(rows), QIndexArg(cols)
        keep += [r, c]
        f = (C.c_int32 * 2)(*((flux, 0) if np.isscalar(flux) else flux))
        return tk_block_matrix(r.c, c.c, f, C.c_void_p(dev.ptr if dev is not None else None))

    def block_gemm_table(self, sym, a_rows, a_cols, a_flux, b_rows, b_cols, b_flux):
        keep = []
        am = self._bm(sym, a_rows, a_cols, a_flux, None, keep)
        bm = self._bm(sym, b_rows, b_cols, b_flux, None, keep)
        n = C.c_int64(0)
        tot = C.c_int64(0)
        lib().tk_block_gemm_table(sym, C.byref(am), C.byref(bm), 0, C.byref(n), None, None, C.byref(tot))
        keys = np.zeros(n.value, dtype=np.uint64)
        offs = np.zeros(n.value, dtype=np.int64)
        _check(lib().tk_block_gemm_table(sym, C.byref(am), C.byref(bm), n.value, C.byref(n),
                                         keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         offs.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(tot)))
        return keys, offs, tot.value

    def block_gemm(self, sym, a_rows, a_cols, a_flux, a_data, b_rows, b_cols, b_flux, b_data):
        """Returns (keys, offsets, data) of C = A B in FlatBlockStorage order."""
        keys, offs, tot = self.block_gemm_table(sym, a_rows, a_cols, a_flux, b_rows, b_cols, b_flux)
        keep = []
        da, db = self.ctx.to_device(np.asarray(a_data, dtype=np.float64)), self.ctx.to_device(
            np.asarray(b_data, dtype=np.float64))
        dc = self.ctx.empty(max(tot, 1))
        am = self._bm(sym, a_rows, a_cols, a_flux, da, keep)
        bm = self._bm(sym, b_rows, b_cols, b_flux, db, keep)
        _check(self.ctx.L.tk_block_gemm(self.ctx.h, sym, C.byref(am), C.byref(bm), C.c_void_p(dc.ptr)), self.ctx.h)
        out = dc.get()[:tot]
        for x in (da, db, dc):
            x.free()
        return keys, offs, out

    def spmv(self, sym, rows, cols, flux, a_data, x):
        keep = []
        da, dx = self.ctx.to_device(np.asarray(a_data, dtype=np.float64)), self.ctx.to_device(
            np.asarray(x, dtype=np.float64))
        n_rows = sum(d for _, d in rows)
        dy = self.ctx.empty(n_rows)
        am = self._bm(sym, rows, cols, flux, da, keep)
        _check(self.ctx.L.tk_spmv(self.ctx.h, sym, C.byref(am), C.c_void_p(dx.ptr), C.c_void_p(dy.ptr)), self.ctx.h)
        out = dy.get()
        for v in (da, dx, dy):
            v.free()
        return out

    def svd_truncated(self, a, max_rank, cutoff):
        """(U, s, Vt, truncation_error) of a row-major matrix (SvdResult, results.rs:10-24)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        m, n = a.shape
        r = min(m, n)
        da = self.ctx.to_device(a)
        du, dv = self.ctx.empty(m * r), self.ctx.empty(r * n)
        s = np.zeros(r)
        rank = C.c_int64(0)
        te = C.c_double(0)
        _check(self.ctx.L.tk_svd_truncated(self.ctx.h, m, n, C.c_void_p(da.ptr), int(max_rank), cutoff,
                                           C.c_void_p(du.ptr), s.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.c_void_p(dv.ptr), C.byref(rank), C.byref(te)), self.ctx.h)
        k = rank.value
        u = du.get()[:m * k].reshape(m, k)
        vt = dv.get()[:k * n].reshape(k, n)
        for x in (da, du, dv):
            x.free()
        return u, s[:k], vt, te.value

    def truncate_svd(self, a, max_bond_dim=200, svd_cutoff=1e-12, min_bond_dim=1):
        """truncate_svd (tk-dmrg/src/truncation/mod.rs:56-127) with a
        TruncationConfig: (U, s, Vt, bond_dim_new, n_discarded, relative
        truncation error)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        m, n = a.shape
        r = min(m, n)
        da = self.ctx.to_device(a)
        du, dv = self.ctx.empty(m * r), self.ctx.empty(r * n)
        s = np.zeros(r)
        bd, nd, te = C.c_int64(0), C.c_int64(0), C.c_double(0)
        _check(self.ctx.L.tk_truncate_svd(self.ctx.h, m, n, C.c_void_p(da.ptr), int(max_bond_dim), float(svd_cutoff),
                                          int(min_bond_dim), C.c_void_p(du.ptr), _dp(s), C.c_void_p(dv.ptr),
                                          C.byref(bd), C.byref(nd), C.byref(te)), self.ctx.h)
        k = bd.value
        u = du.get()[:m * k].reshape(m, k)
        vt = dv.get()[:k * n].reshape(k, n)
        for x in (da, du, dv):
            x.free()
        return u, s[:k], vt, k, nd.value, te.value

    def block_svd_truncate(self, blocks, max_bond, cutoff, min_bond=1, degeneracy_tol=1e-12, sigma2_floor=1e-26):
        """The two-site split's block-diagonal truncated SVD: per block (U_b,
        sigma_b, C_b = U_b^T X_b) and (n_kept, n_discarded, relative error)."""
        blocks = [np.ascontiguousarray(b, dtype=np.float64) for b in blocks]
        rows = np.array([b.shape[0] for b in blocks], dtype=np.int64)
        cols = np.array([b.shape[1] for b in blocks], dtype=np.int64)
        flat = np.concatenate([b.ravel() for b in blocks]) if blocks else np.zeros(1)
        cap_s = int(sum(min(r, c) for r, c in zip(rows, cols)))
        cap_u = int(sum(r * min(r, c) for r, c in zip(rows, cols)))
        cap_c = int(sum(c * min(r, c) for r, c in zip(rows, cols)))
        dx = self.ctx.to_device(flat)
        du, dc = self.ctx.empty(max(cap_u, 1)), self.ctx.empty(max(cap_c, 1))
        keep = np.zeros(max(1, len(blocks)), dtype=np.int64)
        sig = np.zeros(max(1, cap_s))
        nk, nd, te = C.c_int64(0), C.c_int64(0), C.c_double(0)
        _check(self.ctx.L.tk_block_svd_truncate(self.ctx.h, len(blocks), _ip(rows), _ip(cols), C.c_void_p(dx.ptr),
                                                int(max_bond), float(cutoff), int(min_bond), float(degeneracy_tol),
                                                float(sigma2_floor), _ip(keep), _dp(sig), cap_s, C.c_void_p(du.ptr),
                                                cap_u, C.c_void_p(dc.ptr), cap_c, C.byref(nk), C.byref(nd),
                                                C.byref(te)), self.ctx.h)
        uh, ch = du.get(), dc.get()
        out, so, uo, co = [], 0, 0, 0
        for b in range(len(blocks)):
            k, m, n = int(keep[b]), int(rows[b]), int(cols[b])
            out.append((uh[uo:uo + m * k].reshape(m, k), sig[so:so + k].copy(), ch[co:co + k * n].reshape(k, n)))
            so, uo, co = so + k, uo + m * k, co + k * n
        for x in (dx, du, dc):
            x.free()
        return out, nk.value, nd.value, te.value

    def eigh_lowest(self, a, k):
        """The k lowest eigenpairs (eigh_lowest, traits.rs:134-138): (values, n x k vectors)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        da = self.ctx.to_device(a)
        dv = self.ctx.empty(max(1, n * k))
        w = np.zeros(max(1, k))
        _check(self.ctx.L.tk_eigh_lowest(self.ctx.h, n, C.c_void_p(da.ptr), int(k), _dp(w), C.c_void_p(dv.ptr)),
               self.ctx.h)
        v = dv.get()[:n * k].reshape(n, k)
        da.free()
        dv.free()
        return w[:k], v

    def qr(self, a):
        """Thin QR (qr, traits.rs:148): (Q m x k, R k x n), k = min(m, n)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        m, n = a.shape
        k = min(m, n)
        da = self.ctx.to_device(a)
        dq, dr = self.ctx.empty(m * k), self.ctx.empty(k * n)
        _check(self.ctx.L.tk_qr(self.ctx.h, m, n, C.c_void_p(da.ptr), C.c_void_p(dq.ptr), C.c_void_p(dr.ptr)),
               self.ctx.h)
        q, r = dq.get().reshape(m, k), dr.get().reshape(k, n)
        for x in (da, dq, dr):
            x.free()
        return q, r

    def regularized_svd_inverse(self, s, u, vt, delta):
        """V diag(s / (s^2 + delta^2)) U^T (regularized_svd_inverse, traits.rs:150-183)."""
        s = np.ascontiguousarray(s, dtype=np.float64)
        u, vt = np.ascontiguousarray(u, dtype=np.float64), np.ascontiguousarray(vt, dtype=np.float64)
        m, k = u.shape
        n = vt.shape[1]
        du, dv, do = self.ctx.to_device(u), self.ctx.to_device(vt), self.ctx.empty(n * m)
        _check(self.ctx.L.tk_regularized_svd_inverse(self.ctx.h, m, n, k, _dp(s), C.c_void_p(du.ptr),
                                                     C.c_void_p(dv.ptr), float(delta), C.c_void_p(do.ptr)),
               self.ctx.h)
        out = do.get().reshape(n, m)
        for x in (du, dv, do):
            x.free()
        return out

    def eigh(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        da = self.ctx.to_device(a)
        dv = self.ctx.empty(n * n)
        w = np.zeros(n)
        _check(self.ctx.L.tk_eigh(self.ctx.h, n, C.c_void_p(da.ptr), w.ctypes.data_as(C.POINTER(C.c_double)),
                                  C.c_void_p(dv.ptr)), self.ctx.h)
        v = dv.get().reshape(n, n)
        da.free()
        dv.free()
        return w, v

    def davidson(self, matvec_dev, n, x0, diag=None, tol=1e-10, max_subspace=60, max_iter=1000):
        """Lowest eigenpair of a device matvec(x_ptr, y_ptr) (IterativeEigensolver)."""
        cb = MATVEC_FN(lambda x, y, u: matvec_dev(x, y))
        dx = self.ctx.to_device(np.asarray(x0, dtype=np.float64))
        dd = self.ctx.to_device(np.asarray(diag, dtype=np.float64)) if diag is not None else None
        ev, rn = C.c_double(0), C.c_double(0)
        nmv = C.c_int32(0)
        _check(self.ctx.L.tk_davidson(self.ctx.h, n, cb, None, C.c_void_p(dd.ptr) if dd else None, tol, max_subspace,
                                      max_iter, C.c_void_p(dx.ptr), C.byref(ev), C.byref(rn), C.byref(nmv)),
               self.ctx.h)
        x = dx.get()
        dx.free()
        if dd:
            dd.free()
        return ev.value, x, rn.value, nmv.value

    def exp_krylov(self, matvec_dev, n, v, alpha, krylov_dim=30, tol=1e-12, adaptive_tol=0.0):
        """exp(alpha A) v over a device matvec (exp_krylov_f64, tdvp/mod.rs:197-272;
        adaptive_tol > 0: early stop, not in the reference); returns (result,
        Krylov vectors used)."""
        cb = MATVEC_FN(lambda x, y, u: matvec_dev(x, y))
        dv = self.ctx.to_device(np.asarray(v, dtype=np.float64))
        do = self.ctx.empty(n)
        used = C.c_int32(0)
        _check(self.ctx.L.tk_exp_krylov(self.ctx.h, n, cb, None, C.c_void_p(dv.ptr), float(alpha), int(krylov_dim),
                                        float(tol), float(adaptive_tol), C.c_void_p(do.ptr), C.byref(used)),
               self.ctx.h)
        out = do.get()
        dv.free()
        do.free()
        return out, used.value

    def lanczos(self, matvec_dev, n, x0, tol=1e-10, max_krylov=100, max_restarts=1000):
        """Lowest eigenpair by Lanczos with full reorthogonalisation (LanczosSolver)."""
        cb = MATVEC_FN(lambda x, y, u: matvec_dev(x, y))
        dx = self.ctx.to_device(np.asarray(x0, dtype=np.float64))
        ev, rn = C.c_double(0), C.c_double(0)
        nmv = C.c_int32(0)
        _check(self.ctx.L.tk_lanczos(self.ctx.h, n, cb, None, tol, max_krylov, max_restarts, C.c_void_p(dx.ptr),
                                     C.byref(ev), C.byref(rn), C.byref(nmv)), self.ctx.h)
        x = dx.get()
        dx.free()
        return ev.value, x, rn.value, nmv.value


# ---------------------------------------------------------------- H_eff plan
class HeffPlan:
    """tk_heff_plan: H_eff*psi for one bond shape (build_heff_two_site)."""

    def __init__(self, ctx, prec, sym, bond_l, w_l, L_dev, bond_r, w_r, R_dev, mpo_i, mpo_j):
        self.ctx = ctx
        self._keep = []
        ql, wl, qr, wr = QIndexArg(bond_l), QIndexArg(w_l), QIndexArg(bond_r), QIndexArg(w_r)
        self._keep += [ql, wl, qr, wr]
        el = tk_env(ql.c, wl.c, C.c_void_p(L_dev.ptr))
        er = tk_env(qr.c, wr.c, C.c_void_p(R_dev.ptr))
        mi, mj = self._mpo(mpo_i), self._mpo(mpo_j)
        h = C.c_void_p()
        _check(ctx.L.tk_heff_plan_create(ctx.h, PREC[prec], sym, C.byref(el), C.byref(er), C.byref(mi),
                                         C.byref(mj), C.byref(h)), ctx.h)
        self.h = h
        ctx.adopt(self)

    def _mpo(self, m):
        p, a, b = QIndexArg(m.phys), QIndexArg(m.wl), QIndexArg(m.wr)
        w = np.ascontiguousarray(m.w, dtype=np.float64)
        self._keep += [p, a, b, w]
        return tk_mpo_site(p.c, a.c, b.c, w.ctypes.data_as(C.POINTER(C.c_double)))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.tk_heff_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def psi_table(self):
        n = C.c_int64(0)
        tot = C.c_int64(0)
        self.ctx.L.tk_heff_psi_table(self.h, 0, C.byref(n), None, None, C.byref(tot))
        keys = np.zeros(n.value, dtype=np.uint64)
        offs = np.zeros(n.value, dtype=np.int64)
        _check(self.ctx.L.tk_heff_psi_table(self.h, n.value, C.byref(n), keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                                            offs.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(tot)), self.ctx.h)
        return keys, offs, tot.value

    def apply(self, x_dev, y_dev):
        _check(self.ctx.L.tk_heff_apply(self.h, C.c_void_p(x_dev.ptr), C.c_void_p(y_dev.ptr)), self.ctx.h)

    def apply_graph(self, x_dev, y_dev):
        """y = H_eff x replayed from the plan's HIP graph (tk_heff_apply_graph)."""
        _check(self.ctx.L.tk_heff_apply_graph(self.h, C.c_void_p(x_dev.ptr), C.c_void_p(y_dev.ptr)), self.ctx.h)

    def tiling(self, stage):
        """(wg_kind, tile_min) stage 1 or 3 runs with (tk_heff_tiling)."""
        k, t = C.c_int32(0), C.c_int32(0)
        _check(self.ctx.L.tk_heff_tiling(self.h, stage, C.byref(k), C.byref(t)), self.ctx.h)
        return k.value, t.value

    def apply_fp64(self, x_dev, y_dev):
        _check(self.ctx.L.tk_heff_apply_fp64(self.h, C.c_void_p(x_dev.ptr), C.c_void_p(y_dev.ptr)), self.ctx.h)

    def diag(self, d_dev):
        _check(self.ctx.L.tk_heff_diag(self.h, C.c_void_p(d_dev.ptr)), self.ctx.h)

    def stats(self):
        f = C.c_double(0)
        b = C.c_double(0)
        cnt = np.zeros(8, dtype=np.int64)
        _check(self.ctx.L.tk_heff_stats(self.h, C.byref(f), C.byref(b), cnt.ctypes.data_as(C.POINTER(C.c_int64))),
               self.ctx.h)
        return {"flops": f.value, "bytes": b.value, "psi_blocks": int(cnt[0]), "dim": int(cnt[1]),
                "t1_elems": int(cnt[2]), "stage1_items": int(cnt[3]), "stage3_items": int(cnt[4]),
                "stage3_terms": int(cnt[5])}

    def stage_stats(self):
        out = np.zeros(4)
        _check(self.ctx.L.tk_heff_stage_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), self.ctx.h)
        return {"s1_flops": out[0], "s1_bytes": out[1], "s3_flops": out[2], "s3_bytes": out[3]}

    def time(self, x_dev, y_dev, reps):
        a, s1, s3 = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(self.ctx.L.tk_heff_time(self.h, C.c_void_p(x_dev.ptr), C.c_void_p(y_dev.ptr), int(reps), C.byref(a),
                                       C.byref(s1), C.byref(s3)), self.ctx.h)
        return a.value, s1.value, s3.value


# ---------------------------------------------------------------- MPO builders
@dataclass
class MpoSite:
    """Dense W[s_in][s_out][w_l][w_r] with the caller's QIndex sector order."""
    phys: list
    wl: list
    wr: list
    w: np.ndarray


def _op(mat):
    return np.array(mat, dtype=np.float64).reshape(len(mat) if np.ndim(mat) == 2 else int(math.isqrt(len(mat))), -1)


# Fermion basis |0>, |up>, |dn>, |updn> (tk-dsl/src/operators.rs:88, 117-200);
# FermionOp matrices omit every Jordan-Wigner string: c_dn|updn> = +|up>.
F_CDAG_UP = _op([[0, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0]])
F_C_UP = F_CDAG_UP.T.copy()
F_CDAG_DN = _op([[0, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]])
F_C_DN = F_CDAG_DN.T.copy()
F_ID = np.eye(4)
F_PARITY = np.diag([1.0, -1.0, -1.0, 1.0])   # (-1)^(n_up + n_dn)
F_UP_PARITY = np.diag([1.0, -1.0, 1.0, -1.0])  # (-1)^n_up, the on-site string of c_dn
F_NUPNDN = np.diag([0.0, 0.0, 0.0, 1.0])
# charges (N_up, N_dn) of the four states
HUBBARD_PHYS = [((0, 0), 1), ((1, 0), 1), ((0, 1), 1), ((1, 1), 1)]


def hubbard_w(t=1.0, U=4.0, position="bulk"):
    """Hubbard-chain MPO site tensor, bond dim 6, U(1)xU(1) (N_up, N_dn).

    H = -t sum_{i,s} (c+_{i s} c_{i+1 s} + h.c.) + U sum_i n_{i up} n_{i dn}.
    With on-site c_dn = F_up * c_dn(ref) and the inter-site string F, the
    hopping c+_i c_{i+1} = (c+_i F_i) (x) c_{i+1} and c+_{i+1} c_i = (F_i c_i) (x) c+_{i+1}.
    MPO states (caller order, grouped by charge):
      0 start (0,0), 1 done (0,0), 2 after c+_up F (-1,0), 3 after F c_up (+1,0),
      4 after c+_dn F (0,-1), 5 after F c_dn (0,+1).
    """
    cu, cdu = F_C_UP, F_CDAG_UP
    cd = F_UP_PARITY @ F_C_DN
    cdd = cd.T.copy()
    ops = {}
    ops[(0, 0)] = F_ID
    ops[(1, 1)] = F_ID
    ops[(0, 1)] = U * F_NUPNDN
    ops[(0, 2)] = cdu @ F_PARITY
    ops[(0, 3)] = F_PARITY @ cu
    ops[(0, 4)] = cdd @ F_PARITY
    ops[(0, 5)] = F_PARITY @ cd
    ops[(2, 1)] = -t * cu
    ops[(3, 1)] = -t * cdu
    ops[(4, 1)] = -t * cd
    ops[(5, 1)] = -t * cdd
    wq = [((0, 0), 2), ((-1, 0), 1), ((1, 0), 1), ((0, -1), 1), ((0, 1), 1)]
    return _assemble(ops, 4, 6, HUBBARD_PHYS, wq, position, start=0, done=1)


# Spin-1/2 basis (|up>, |dn>), S+ = [[0,1],[0,0]] (operators.rs:31-58); U(1) charge 2*Sz.
S_PLUS = _op([[0, 1], [0, 0]])
S_MINUS = _op([[0, 0], [1, 0]])
S_Z = _op([[0.5, 0], [0, -0.5]])
S_ID = np.eye(2)
SPIN_PHYS = [((1, 0), 1), ((-1, 0), 1)]


def heisenberg_w(J=1.0, position="bulk"):
    """Heisenberg-chain MPO site tensor, bond dim 5, U(1) 2*Sz.

    H = J sum_i (S+_i S-_{i+1}/2 + S-_i S+_{i+1}/2 + Sz_i Sz_{i+1}).
    States: 0 start, 1 done, 2 after Sz (all charge 0), 3 after S+ (-2), 4 after S- (+2).
    """
    ops = {(0, 0): S_ID, (1, 1): S_ID, (0, 2): S_Z, (0, 3): S_PLUS, (0, 4): S_MINUS,
           (2, 1): J * S_Z, (3, 1): 0.5 * J * S_MINUS, (4, 1): 0.5 * J * S_PLUS}
    wq = [((0, 0), 3), ((-2, 0), 1), ((2, 0), 1)]
    return _assemble(ops, 2, 5, SPIN_PHYS, wq, position, start=0, done=1)


# Real-time evolution in real arithmetic.  A complex state psi = psi_R + i psi_I
# is stored as the real state Phi = |R> psi_R + |I> psi_I with one extra
# charge-0 site of two states (the "ancilla") in front of the chain.  With
# J|R> = -|I>, J|I> = |R>, exp(t J (x) H) Phi carries exp(-i H t) psi, so the
# two-site TDVP step with tau = -t evolves in real time; J (x) H is real
# antisymmetric and the Krylov exponential (Arnoldi) needs no symmetry.
ANCILLA_PHYS = [((0, 0), 2)]


def complex_time_mpo(mpo):
    """MPO of J (x) H on [ancilla] + the sites of `mpo` (W[s_in][s_out] = <s_out|J|s_in>)."""
    w = np.zeros((2, 2, 1, 1))
    w[0, 1, 0, 0] = -1.0  # <I|J|R>
    w[1, 0, 0, 0] = 1.0   # <R|J|I>
    return [MpoSite(ANCILLA_PHYS, list(mpo[0].wl), list(mpo[0].wl), w)] + list(mpo)


def complex_state(bonds, sites, part="R"):
    """(bonds, sites) of |part> (x) phi on [ancilla] + sites for a real state phi."""
    anc = np.array([1.0, 0.0]) if part == "R" else np.array([0.0, 1.0])
    return [bonds[0], bonds[0]] + list(bonds[1:]), [anc] + [np.asarray(x, dtype=np.float64) for x in sites]


def greens_function_real_time(ctx, mpo, sym, bonds, sites, e0, dt, n_steps, max_bond=256, krylov_dim=30,
                              krylov_tol=1e-12, svd_cutoff=1e-12, start=None, adaptive_tol=0.0):
    """G(t_k) = -i exp(i e0 t_k) <phi| exp(-i H t_k) |phi>, t_k = k dt, k = 0..n_steps,
    for the real normalised state phi = (bonds, sites) on the sites of `mpo`
    (for example O psi_0 from Dmrg.apply_site_op, then multiply by its norm^2).

    Steps 3-4 of compute_greens_function_tdvp (tk-dmft/src/spectral/tdvp.rs:55-108):
    two-site TDVP in real time on the ancilla form (tau = -dt per step) and the
    overlaps <R phi|Phi(t)>, <I phi|Phi(t)> (mps_overlap).  `start` replaces the
    initial complex state |R> phi (bonds, sites on [ancilla] + sites)."""
    cmpo = complex_time_mpo(mpo)
    ket = Dmrg(ctx, cmpo, sym, schedule=(max_bond,), svd_cutoff=svd_cutoff)
    ket.set_state(*(start if start is not None else complex_state(bonds, sites, "R")))
    ket.prepare()
    bra_r = Dmrg(ctx, cmpo, sym, schedule=(max_bond,))
    bra_r.set_state(*complex_state(bonds, sites, "R"))
    bra_i = Dmrg(ctx, cmpo, sym, schedule=(max_bond,))
    bra_i.set_state(*complex_state(bonds, sites, "I"))
    ts, g, acc = [], [], 0.0
    for k in range(n_steps + 1):
        if k > 0:
            acc += ket.tdvp_step(-dt, max_bond, krylov_dim, krylov_tol, adaptive_tol)["log_norm"]
        t = k * dt
        amp = (bra_r.overlap(ket) + 1j * bra_i.overlap(ket)) * math.exp(acc)
        ts.append(t)
        g.append(-1j * complex(math.cos(e0 * t), math.sin(e0 * t)) * amp)
    for e in (ket, bra_r, bra_i):
        e.close()
    return np.array(ts), np.array(g)


def snake_path(lx, ly):
    """Boustrophedon order of an lx x ly grid (tk-dsl/src/lattice/square.rs:64-80):
    row y = 0, 1, ... of lx sites each (site y * lx + x), even rows left to
    right, odd rows right to left."""
    out = []
    for y in range(ly):
        xs = range(lx) if y % 2 == 0 else range(lx - 1, -1, -1)
        out += [y * lx + x for x in xs]
    return out


def cylinder_bonds(width, length):
    """Nearest-neighbour bonds of a width x length cylinder, as chain positions
    of the snake order: the open square lattice's bonds
    (tk-dsl/src/lattice/square.rs:16-33, rows of `width` sites, `length` rows)
    plus the wrap bond (0, width - 1) of every row when width > 2."""
    pos = {site: k for k, site in enumerate(snake_path(width, length))}
    bonds = []
    for y in range(length):
        for x in range(width):
            s = y * width + x
            if x + 1 < width:
                bonds.append((s, s + 1))
            if y + 1 < length:
                bonds.append((s, s + width))
        if width > 2:
            bonds.append((y * width, y * width + width - 1))
    return sorted(tuple(sorted((pos[a], pos[b]))) for a, b in bonds)


def bond_heisenberg_mpo(n, bonds, J=1.0):
    """Exact finite-state-automaton MPO of H = J sum_{(i,j) in bonds} S_i.S_j
    on n chain positions, U(1) 2*Sz.  The states on the cut after site c are
    start, done and, for every source i <= c with a partner j > c (ascending
    i), "Sz_i / S+_i / S-_i placed": bond dim 2 + 3 * (open sources).
    Caller order on each cut: start, done, Sz_i... (charge 0); S+_i... (-2);
    S-_i... (+2).  Returns one MpoSite per site (first: left row `start`,
    last: right column `done`)."""
    partners = [set() for _ in range(n)]
    for i, j in bonds:
        if not (0 <= i < j < n):
            raise ValueError(f"bond {(i, j)} is not an ordered pair of positions below {n}")
        partners[i].add(j)
    # sources open on the cut after site c
    cuts = [[i for i in range(c + 1) if any(j > c for j in partners[i])] for c in range(n)]
    cuts = [[]] + cuts  # the cut before site 0

    def states(src):
        k = len(src)
        idx = {("start",): 0, ("done",): 1}
        for t, i in enumerate(src):
            idx[("z", i)] = 2 + t
            idx[("p", i)] = 2 + k + t
            idx[("m", i)] = 2 + 2 * k + t
        q = [((0, 0), 2 + k)] + ([((-2, 0), k), ((2, 0), k)] if k else [])
        return idx, 2 + 3 * k, q

    out = []
    for s in range(n):
        li, lw, lq = states(cuts[s])
        ri, rw, rq = states(cuts[s + 1])
        ops = {(0, 0): S_ID, (1, 1): S_ID}
        for i in cuts[s]:
            if i in cuts[s + 1]:
                for t in "zpm":
                    ops[(li[(t, i)], ri[(t, i)])] = S_ID
            if s in partners[i]:
                ops[(li[("z", i)], 1)] = J * S_Z
                ops[(li[("p", i)], 1)] = 0.5 * J * S_MINUS
                ops[(li[("m", i)], 1)] = 0.5 * J * S_PLUS
        if s in cuts[s + 1]:
            ops[(0, ri[("z", s)])] = S_Z
            ops[(0, ri[("p", s)])] = S_PLUS
            ops[(0, ri[("m", s)])] = S_MINUS
        full = np.zeros((2, 2, lw, rw))
        for (a, b), O in ops.items():
            full[:, :, a, b] = O.T
        if n == 1:
            out.append(MpoSite(SPIN_PHYS, [((0, 0), 1)], [((0, 0), 1)], full[:, :, 0:1, 1:2].copy()))
        elif s == 0:
            out.append(MpoSite(SPIN_PHYS, [((0, 0), 1)], rq, full[:, :, 0:1, :].copy()))
        elif s == n - 1:
            out.append(MpoSite(SPIN_PHYS, lq, [((0, 0), 1)], full[:, :, :, 1:2].copy()))
        else:
            out.append(MpoSite(SPIN_PHYS, lq, rq, full))
    return out


def cylinder_mpo(width=6, length=12, J=1.0):
    """J1 Heisenberg on a width x length cylinder in snake order (72 sites and
    MPO bond dims up to 23 for 6 x 12)."""
    return bond_heisenberg_mpo(width * length, cylinder_bonds(width, length), J)


# ---------------------------------------------------------------- OpSum and its MPO
@dataclass
class SiteType:
    """Local basis of a site: QIndex of the physical leg, named operator
    matrices (row = out, column = in), the fermionic ones, and the string
    operator carried between the two ends of a fermionic pair."""
    phys: list
    ops: dict
    fermionic: frozenset
    string: np.ndarray


def _spin_half():
    return SiteType(SPIN_PHYS, {"Id": S_ID, "Sz": S_Z, "S+": S_PLUS, "S-": S_MINUS}, frozenset(), S_ID)


def _electron():
    """Fermion basis |0>, |up>, |dn>, |updn> (tk-dsl/src/operators.rs:88, 117-200);
    the down operators carry the on-site string (-1)^n_up."""
    cu, cdu = F_C_UP, F_CDAG_UP
    cd = F_UP_PARITY @ F_C_DN
    cdd = cd.T.copy()
    nu, nd = cdu @ cu, cdd @ cd
    ops = {"Id": F_ID, "F": F_PARITY, "Cup": cu, "Cdagup": cdu, "Cdn": cd, "Cdagdn": cdd, "Nup": nu, "Ndn": nd,
           "Ntot": nu + nd, "Nupdn": F_NUPNDN, "Sz": 0.5 * (nu - nd), "S+": cdu @ cd, "S-": cdd @ cu}
    return SiteType(HUBBARD_PHYS, ops, frozenset({"Cup", "Cdagup", "Cdn", "Cdagdn"}), F_PARITY)


SITE_TYPES = {"spin-1/2": _spin_half, "electron": _electron}


class OpSum:
    """Runtime accumulator of weighted operator products
    (tk-dsl/src/opsum.rs:51-143): terms (coefficient, [(operator, site), ...])
    in operator order.  `compile_mpo` turns sums of one- and two-site terms
    into an exact finite-state MPO."""

    def __init__(self, n_sites=None):
        self.n_sites = n_sites
        self.terms = []

    def push_term(self, coeff, *factors):
        """Add coeff * O1(site1) O2(site2) ... (factors (name, site))."""
        for name, site in factors:
            if self.n_sites is not None and not (0 <= site < self.n_sites):
                raise ValueError(f"site {site} outside a lattice of {self.n_sites} sites")
        self.terms.append((float(coeff), [(str(n), int(i)) for n, i in factors]))
        return self

    add = push_term

    def n_terms(self):
        return len(self.terms)

    def iter_terms(self):
        return iter(self.terms)

    def scale(self, factor):
        self.terms = [(c * factor, f) for c, f in self.terms]

    def extend(self, other):
        self.terms.extend(other.terms)


def _flux(mat, phys):
    """q(out) - q(in) of every nonzero element of a local operator (one value)."""
    qs = [(_qkey(q)) for q, d in phys for _ in range(d)]
    fl = {(qs[o][0] - qs[i][0], qs[o][1] - qs[i][1]) for o, i in zip(*np.nonzero(np.abs(mat) > 0))}
    if len(fl) > 1:
        raise ValueError("operator does not carry a definite charge")
    return fl.pop() if fl else (0, 0)


def compile_mpo(opsum, n_sites, site_type="spin-1/2"):
    """Exact finite-state MPO of an OpSum of one- and two-site terms (the
    reference's compile step, tk-dmrg/src/mpo/mod.rs:96-162, is a stub).
    Factors on one site are multiplied in operator order.  A two-site term
    c A_a B_b becomes a channel opened at min(a, b) and closed at max(a, b);
    fermionic pairs are reordered to the left site first (one sign) and
    carry the string between the ends (left factor A F), as in the Hubbard
    MPO.  Channels with the same left site and left matrix share one state,
    so the MPO bond on a cut is 2 + (open channels).  Returns one MpoSite
    per site (first: left row `start`, last: right column `done`)."""
    st = SITE_TYPES[site_type]() if isinstance(site_type, str) else site_type
    d = len(st.string)
    onsite = [np.zeros((d, d)) for _ in range(n_sites)]
    chans = {}  # (left site, key) -> [left matrix, fermionic, {right site: right matrix sum}]
    for coeff, factors in opsum.iter_terms():
        by_site = {}
        order = []
        for name, site in factors:
            if name not in st.ops:
                raise ValueError(f"unknown operator {name}")
            if site in by_site:
                if name in st.fermionic:
                    raise ValueError("fermionic operators on one site must be adjacent and are not supported")
                by_site[site] = by_site[site] @ st.ops[name]
            else:
                by_site[site] = st.ops[name]
                order.append((site, name in st.fermionic))
        if len(order) == 1:
            onsite[order[0][0]] += coeff * by_site[order[0][0]]
            continue
        if len(order) != 2:
            raise ValueError("only one- and two-site terms compile")
        (a, fa), (b, fb) = order
        if fa != fb:
            raise ValueError("a term with one fermionic factor changes the parity")
        sign = 1.0
        if a > b:
            a, b = b, a
            if fa:
                sign = -1.0  # B_b A_a = -A_a B_b for two fermion operators
        left = by_site[a] @ st.string if fa else by_site[a]
        key = (a, left.tobytes(), fa)
        ch = chans.setdefault(key, [left, fa, {}])
        ch[2][b] = ch[2].get(b, 0.0) + sign * coeff * by_site[b]
    # open channels on each cut (the cut after site c), in a fixed order
    keys = sorted(chans, key=lambda k: (k[0], k[1], k[2]))
    flux = {k: _flux(chans[k][0], st.phys) for k in keys}
    cuts = [[]]
    for c in range(n_sites):
        cuts.append([k for k in keys if k[0] <= c and max(chans[k][2]) > c])

    def states(open_keys):
        chg = sorted({(-flux[k][0], -flux[k][1]) for k in open_keys}, key=lambda q: (q != (0, 0), q))
        idx = {"start": 0, "done": 1}
        q = [((0, 0), 2)]
        nxt = 2
        for qq in chg:
            ks = [k for k in open_keys if (-flux[k][0], -flux[k][1]) == qq]
            for k in ks:
                idx[k] = nxt
                nxt += 1
            if qq == (0, 0):
                q[0] = ((0, 0), 2 + len(ks))
            else:
                q.append((qq, len(ks)))
        return idx, nxt, q

    out = []
    eye = np.eye(d)
    for s in range(n_sites):
        li, lw, lq = states(cuts[s])
        ri, rw, rq = states(cuts[s + 1])
        full = np.zeros((d, d, lw, rw))

        def put(a, b, O):
            full[:, :, a, b] += O.T

        put(0, 0, eye)
        put(1, 1, eye)
        if np.any(onsite[s]):
            put(0, 1, onsite[s])
        for k in cuts[s]:
            left, fa, closes = chans[k]
            if k in ri:
                put(li[k], ri[k], st.string if fa else eye)
            if s in closes:
                put(li[k], 1, closes[s])
        for k in cuts[s + 1]:
            if k[0] == s:
                put(0, ri[k], chans[k][0])
        if n_sites == 1:
            out.append(MpoSite(st.phys, [((0, 0), 1)], [((0, 0), 1)], full[:, :, 0:1, 1:2].copy()))
        elif s == 0:
            out.append(MpoSite(st.phys, [((0, 0), 1)], rq, full[:, :, 0:1, :].copy()))
        elif s == n_sites - 1:
            out.append(MpoSite(st.phys, lq, [((0, 0), 1)], full[:, :, :, 1:2].copy()))
        else:
            out.append(MpoSite(st.phys, lq, rq, full))
    return out


def _states(qidx):
    """Charge of every state of a QIndex list, in caller order."""
    return [_qkey(q) for q, d in qidx for _ in range(int(d))]


def _merge_bond(qa, qb):
    """Direct sum of two bond QIndex lists: one sector per charge (first
    appearance order), A's states before B's.  Returns (qindex, ia, ib), the
    new index of every state of A and of B."""
    sa, sb = _states(qa), _states(qb)
    order = []
    for q in sa + sb:
        if q not in order:
            order.append(q)
    ia, ib, qidx, nxt = [0] * len(sa), [0] * len(sb), [], 0
    for q in order:
        start = nxt
        for src, idx in ((sa, ia), (sb, ib)):
            for i, c in enumerate(src):
                if c == q:
                    idx[i] = nxt
                    nxt += 1
        qidx.append((q, nxt - start))
    return qidx, np.array(ia, dtype=np.int64), np.array(ib, dtype=np.int64)


def add_mpo(a, b, alpha=1.0, beta=1.0):
    """MPO of alpha A + beta B on the same sites: the direct sum of the bond
    spaces (block diagonal in the bulk, a row on the first site, a column on
    the last), the sum the compression pass follows in the reference's plan
    (MPO::add_scaled, tk-dmrg/src/mpo/mod.rs:146-162).  The first site
    carries the coefficients."""
    n = len(a)
    if n == 0 or len(b) != n:
        raise ValueError("add_mpo: the two MPOs must have the same nonzero length")
    out = []
    for s in range(n):
        A, B = a[s], b[s]
        if [(_qkey(q), int(d)) for q, d in A.phys] != [(_qkey(q), int(d)) for q, d in B.phys]:
            raise ValueError(f"add_mpo: site {s} has different physical legs")
        if n == 1:
            out.append(MpoSite(A.phys, A.wl, A.wr, alpha * A.w + beta * B.w))
            continue
        d = A.w.shape[0]
        if s == 0:
            wl, la, lb = list(A.wl), np.array([0]), np.array([0])
        else:
            wl, la, lb = _merge_bond(A.wl, B.wl)
        if s == n - 1:
            wr, ra, rb = list(A.wr), np.array([0]), np.array([0])
        else:
            wr, ra, rb = _merge_bond(A.wr, B.wr)
        w = np.zeros((d, d, sum(x for _, x in wl), sum(x for _, x in wr)))
        ca, cb = (alpha, beta) if s == 0 else (1.0, 1.0)
        w[np.ix_(range(d), range(d), la, ra)] += ca * A.w
        w[np.ix_(range(d), range(d), lb, rb)] += cb * B.w
        out.append(MpoSite(A.phys, wl, wr, w))
    return out


def _bond_blocks(W, side):
    """Charge blocks of one site tensor seen as a matrix over its `side` bond.

    side "right": rows (w_l, s_in, s_out), columns w_r, one block per w_r
    charge Q over the rows of charge q(w_l) + q(s_in) - q(s_out) = Q.
    side "left": rows w_l, columns (s_in, s_out, w_r), one block per w_l charge
    Q over the columns of charge q(s_out) + q(w_r) - q(s_in) = Q.
    Returns the matrix and [(Q, bond state indices, other-side indices)]."""
    ql, qr, qp = _states(W.wl), _states(W.wr), _states(W.phys)
    d = len(qp)
    if side == "right":
        M = W.w.transpose(2, 0, 1, 3).reshape(-1, W.w.shape[3])
        other = [(ql[l][0] + qp[i][0] - qp[o][0], ql[l][1] + qp[i][1] - qp[o][1])
                 for l in range(len(ql)) for i in range(d) for o in range(d)]
        bond, qidx = qr, W.wr
    else:
        M = W.w.transpose(2, 0, 1, 3).reshape(W.w.shape[2], -1)
        other = [(qp[o][0] + qr[r][0] - qp[i][0], qp[o][1] + qr[r][1] - qp[i][1])
                 for i in range(d) for o in range(d) for r in range(len(qr))]
        bond, qidx = ql, W.wl
    out, seen = [], set()
    for q, _ in qidx:
        q = _qkey(q)
        if q in seen:
            raise ValueError("compress_mpo: a bond charge appears in two sectors")
        seen.add(q)
        b = np.array([j for j, c in enumerate(bond) if c == q], dtype=np.int64)
        o = np.array([j for j, c in enumerate(other) if c == q], dtype=np.int64)
        if len(b) and len(o):
            out.append((q, b, o))
    # entries outside every block would break charge conservation
    inb = np.zeros(M.shape, dtype=bool)
    for _, b, o in out:
        if side == "right":
            inb[np.ix_(o, b)] = True
        else:
            inb[np.ix_(b, o)] = True
    if np.any(M[~inb] != 0):
        raise ValueError("compress_mpo: the site tensor does not conserve the bond charges")
    return M, out


def compress_mpo(ctx, mpo, max_bond_dim=None, svd_cutoff=1e-12, validate_compression=True, compression_tol=1e-8):
    """SVD compression of an MPO (MpoCompressionConfig and the compress step,
    techspec/6_tech-spec_tk-dmrg.md:380-453, tk-dmrg/src/mpo/mod.rs:96-162,
    where it is a stub).

    A left-to-right pass brings every site but the last to left-orthonormal
    form by charge-block SVDs over its right bond, dropping only directions
    below the split's 1e-26 sigma^2 floor (exact redundancy: a sum of equal
    MPOs, decaying couplings).  A right-to-left pass then splits every bond
    into its charge blocks on the device (HipBackend.block_svd_truncate over
    the transposed blocks), selects the kept values across the bond's blocks
    by the library's truncation rule (truncation_select: at most
    max_bond_dim, sigma > svd_cutoff * sigma_max), keeps V^T on the site and
    moves U S into the left neighbour.  Each right-to-left split happens with
    everything to its left orthonormal, so its sigmas are the operator's
    Schmidt values on that bond and ||H_c - H||_F^2 / ||H||_F^2 is at most
    the sum of the bonds' discarded weights.  The norm, left on the first
    site, is spread evenly over the sites at the end.

    Returns (mpo, info): info["bond_dims"] (n + 1 values), info["discarded"]
    (relative discarded weight per inner bond) and info["error_bound"] =
    sqrt(sum discarded).  With validate_compression an error bound above
    compression_tol raises ValueError (the reference's validation step)."""
    n = len(mpo)
    sites = [MpoSite(list(W.phys), list(W.wl), list(W.wr), np.array(W.w, dtype=np.float64)) for W in mpo]
    if n <= 1:
        return sites, {"bond_dims": [1] * (n + 1), "discarded": [], "error_bound": 0.0}
    be = HipBackend(ctx)
    unlimited = 1 << 40
    # left-to-right: left-orthonormal sites 0..n-2, exact redundancy removed
    for s in range(n - 1):
        W = sites[s]
        M, blocks = _bond_blocks(W, "right")
        mats = [M[np.ix_(rows, cols)] for _, cols, rows in blocks]
        res, nk, _, _ = be.block_svd_truncate(mats, unlimited, 0.0, min_bond=1, degeneracy_tol=0.0)
        k_tot = sum(r[0].shape[1] for r in res)
        Mn = np.zeros((M.shape[0], k_tot))
        T = np.zeros((k_tot, M.shape[1]))
        wr, off = [], 0
        for (q, cols, rows), (u, _, c) in zip(blocks, res):
            k = u.shape[1]
            if k == 0:
                continue
            Mn[np.ix_(rows, range(off, off + k))] = u
            T[np.ix_(range(off, off + k), cols)] = c
            wr.append((q, k))
            off += k
        if k_tot == 0:
            raise ValueError("compress_mpo: the operator is zero")
        d = W.w.shape[0]
        W.w = Mn.reshape(W.w.shape[2], d, d, k_tot).transpose(1, 2, 0, 3).copy()
        W.wr = wr
        nxt = sites[s + 1]
        nxt.w = np.einsum("ab,iobr->ioar", T, nxt.w)
        nxt.wl = list(wr)
    # right-to-left: truncating splits with everything to the left orthonormal
    discarded = [0.0] * (n - 1)
    for s in range(n - 1, 0, -1):
        W = sites[s]
        M, blocks = _bond_blocks(W, "left")
        mats = [M[np.ix_(rows, cols)].T for _, rows, cols in blocks]
        res, _, _, _ = be.block_svd_truncate(mats, unlimited, 0.0, min_bond=1, degeneracy_tol=0.0)
        sig = [r[1] for r in res]
        total = sum(float(np.sum(x * x)) for x in sig)
        keep, _, _, _ = truncation_select(sig, unlimited if max_bond_dim is None else max_bond_dim, svd_cutoff,
                                          min_bond=1)
        discarded[s - 1] = sum(float(np.sum(x[k:] ** 2)) for x, k in zip(sig, keep)) / total if total > 0 else 0.0
        k_tot = int(sum(keep))
        Mn = np.zeros((k_tot, M.shape[1]))
        T = np.zeros((M.shape[0], k_tot))
        wl, off = [], 0
        for (q, rows, cols), (u, _, c), k in zip(blocks, res, keep):
            if k == 0:
                continue
            Mn[np.ix_(range(off, off + k), cols)] = u[:, :k].T
            T[np.ix_(rows, range(off, off + k))] = c[:k].T
            wl.append((q, k))
            off += k
        d = W.w.shape[0]
        W.w = Mn.reshape(k_tot, d, d, W.w.shape[3]).transpose(1, 2, 0, 3).copy()
        W.wl = wl
        prv = sites[s - 1]
        prv.w = np.einsum("iolr,rk->iolk", prv.w, T)
        prv.wr = list(wl)
    norm = float(np.linalg.norm(sites[0].w))
    if norm > 0:
        f = norm ** (1.0 / n)
        sites[0].w *= f / norm
        for W in sites[1:]:
            W.w *= f
    bound = math.sqrt(sum(discarded))
    if validate_compression and bound > compression_tol:
        raise ValueError(f"compress_mpo: error bound {bound:.3e} above compression_tol {compression_tol:.3e}")
    dims = [1] + [sum(int(x) for _, x in W.wr) for W in sites[:-1]] + [1]
    return sites, {"bond_dims": dims, "discarded": discarded, "error_bound": bound}


def hubbard_opsum(bonds, n_sites, t=1.0, U=4.0):
    """-t sum_{(i,j) in bonds, s} (c+_{is} c_{js} + h.c.) + U sum_i n_i,up n_i,dn."""
    os_ = OpSum(n_sites)
    for i, j in bonds:
        for up, cd, c in (("up", "Cdagup", "Cup"), ("dn", "Cdagdn", "Cdn")):
            os_.add(-t, (cd, i), (c, j))
            os_.add(-t, (cd, j), (c, i))
    for i in range(n_sites):
        os_.add(U, ("Nupdn", i))
    return os_


def heisenberg_opsum(bonds, n_sites, J=1.0):
    """J sum_{(i,j) in bonds} S_i . S_j."""
    os_ = OpSum(n_sites)
    for i, j in bonds:
        os_.add(J, ("Sz", i), ("Sz", j))
        os_.add(0.5 * J, ("S+", i), ("S-", j))
        os_.add(0.5 * J, ("S-", i), ("S+", j))
    return os_


def synthetic_u1_bond(total=4096, qmax=30, width=8.0):
    """U(1) (2*Sz) bond of the cylinder microbenchmark: even charges
    -qmax..qmax, dims proportional to exp(-q^2/(2 width^2)), every sector at
    least 1, rounded to `total` by largest remainder."""
    qs = list(range(-qmax, qmax + 1, 2))
    wts = np.array([math.exp(-q * q / (2 * width * width)) for q in qs])
    raw = wts / wts.sum() * (total - len(qs))
    dims = np.floor(raw).astype(int) + 1
    rem = total - dims.sum()
    order = np.argsort(-(raw - np.floor(raw)), kind="stable")
    for i in order[:rem]:
        dims[i] += 1
    return [((q, 0), int(dd)) for q, dd in zip(qs, dims)]


def c4_problem(seed=4096, total=4096, width=6, length=12, site=None):
    """Inputs of the cylinder H_eff*psi microbenchmark: J1 Heisenberg on the
    width x length cylinder (snake order), the two-site bond (site, site + 1)
    (default: the centre), U(1) bonds of `total` states (31 sectors) on both
    sides, L, R and psi N(0,1) from PCG64(seed) in FlatBlockStorage order
    (L, then R, then psi)."""
    mpo = cylinder_mpo(width, length)
    site = len(mpo) // 2 - 1 if site is None else site
    Wi, Wj = mpo[site], mpo[site + 1]
    bl = synthetic_u1_bond(total)
    br = list(bl)
    _, _, _, l_tot = env_table(SYM_U1, bl, Wi.wl)
    _, _, _, r_tot = env_table(SYM_U1, br, Wj.wr)
    _, _, _, p_tot = block_table(SYM_U1, [Wi.phys, Wj.phys, bl, br], [DIR_OUT, DIR_OUT, DIR_OUT, DIR_IN])
    rng = np.random.default_rng(seed)
    L = rng.standard_normal(l_tot)
    R = rng.standard_normal(r_tot)
    x = rng.standard_normal(p_tot)
    return {"bond_l": bl, "bond_r": br, "W": Wi, "Wj": Wj, "L": L, "R": R, "x": x, "sym": SYM_U1, "site": site}


def _assemble(ops, d, w, phys, wq, position, start, done):
    full = np.zeros((d, d, w, w))
    for (a, b), O in ops.items():
        # W[s_in][s_out][a][b] = <s_out| O |s_in> = O[s_out, s_in]
        full[:, :, a, b] = O.T
    if position == "bulk":
        return MpoSite(phys, wq, wq, full)
    if position == "left":
        return MpoSite(phys, [((0, 0), 1)], wq, full[:, :, start:start + 1, :].copy())
    if position == "right":
        return MpoSite(phys, wq, [((0, 0), 1)], full[:, :, :, done:done + 1].copy())
    if position == "single":
        return MpoSite(phys, [((0, 0), 1)], [((0, 0), 1)], full[:, :, start:start + 1, done:done + 1].copy())
    raise ValueError(position)


# ---------------------------------------------------------------- synthetic bond
def synthetic_uu_bond(total=1024, rmax=6, l1max=9, width=2.6):
    """U(1)xU(1) bond of the single H_eff*psi microbenchmark.

    Charges (a, b) with |a|,|b| <= rmax and |a|+|b| <= l1max (145 sectors for
    the defaults); dims proportional to exp(-(a^2+b^2)/(2 width^2)), every
    sector at least 1, rounded to `total` by largest remainder.
    """
    qs = [(a, b) for a in range(-rmax, rmax + 1) for b in range(-rmax, rmax + 1) if abs(a) + abs(b) <= l1max]
    w = np.array([math.exp(-(a * a + b * b) / (2 * width * width)) for a, b in qs])
    n = len(qs)
    raw = (total - n) * w / w.sum()  # one state reserved per sector
    base = np.floor(raw).astype(np.int64)
    rem = int(total - n - base.sum())
    frac = raw - base
    order = sorted(range(n), key=lambda i: (-frac[i], qs[i]))
    for i in order[:rem]:
        base[i] += 1
    dims = base + 1
    assert int(dims.sum()) == total
    return [(q, int(d)) for q, d in zip(qs, dims)]


def _leg_args(legs, dirs, flux):
    qs = [QIndexArg(l) for l in legs]
    arr = (tk_qindex * len(qs))(*[q.c for q in qs])
    d = np.ascontiguousarray(dirs, dtype=np.int32)
    f = np.ascontiguousarray(_qkey(flux), dtype=np.int32)
    return qs, arr, d, f


def block_permute(ctx, sym, legs, dirs, flux, perm, x):
    """BlockSparseTensor::permute on the device; x and the result are host
    arrays in the block-table layouts of legs and of the permuted legs."""
    qs, arr, d, f = _leg_args(legs, dirs, flux)
    p = np.ascontiguousarray(perm, dtype=np.int32)
    _, _, _, tot = block_table(sym, [legs[i] for i in perm], [dirs[i] for i in perm], _qkey(flux))
    dx = ctx.to_device(np.asarray(x, dtype=np.float64))
    dy = ctx.empty(max(1, tot))
    _check(ctx.L.tk_block_permute(ctx.h, sym, len(legs), arr, d.ctypes.data_as(C.POINTER(C.c_int32)),
                                  f.ctypes.data_as(C.POINTER(C.c_int32)), p.ctypes.data_as(C.POINTER(C.c_int32)),
                                  C.c_void_p(dx.ptr), len(x), C.c_void_p(dy.ptr), tot), ctx.h)
    y = dy.get()[:tot]
    dx.free()
    dy.free()
    return y


def _fuse_call(fn, ctx, sym, legs, dirs, flux, start, end, x, n_out):
    qs, arr, d, f = _leg_args(legs, dirs, flux)
    dx = ctx.to_device(np.asarray(x, dtype=np.float64))
    dy = ctx.empty(max(1, n_out))
    _check(fn(ctx.h, sym, len(legs), arr, d.ctypes.data_as(C.POINTER(C.c_int32)),
              f.ctypes.data_as(C.POINTER(C.c_int32)), start, end, C.c_void_p(dx.ptr), len(x), C.c_void_p(dy.ptr),
              n_out), ctx.h)
    y = dy.get()[:n_out]
    dx.free()
    dy.free()
    return y


def block_fuse(ctx, sym, legs, dirs, flux, start, end, x):
    """BlockSparseTensor::fuse_legs(start..end) on the device (host arrays)."""
    return _fuse_call(ctx.L.tk_block_fuse, ctx, sym, legs, dirs, flux, start, end, x, len(x))


def block_split(ctx, sym, legs, dirs, flux, start, end, xf):
    """BlockSparseTensor::split_leg on the device: xf in the fused layout of
    (legs, start, end); the result in the layout of legs."""
    return _fuse_call(ctx.L.tk_block_split, ctx, sym, legs, dirs, flux, start, end, xf, len(xf))


def env_table(sym, bond, wq):
    """Block table of an environment (bond, w, bond), q(bra) = q(ket) - q(w)."""
    return block_table(sym, [bond, wq, bond], [DIR_IN, DIR_IN, DIR_OUT])


def heff_problem(bond_l, bond_r, W, seed=0, sym=SYM_U1U1):
    """Two-site H_eff*psi inputs on given bonds (same MPO site on both sides):
    L, R and psi block entries N(0,1) from PCG64(`seed`) in FlatBlockStorage
    order (L, then R, then psi)."""
    _, _, _, l_tot = env_table(sym, bond_l, W.wl)
    _, _, _, r_tot = env_table(sym, bond_r, W.wr)
    _, _, _, p_tot = block_table(sym, [W.phys, W.phys, bond_l, bond_r], [DIR_OUT, DIR_OUT, DIR_OUT, DIR_IN])
    rng = np.random.default_rng(seed)
    L = rng.standard_normal(l_tot)
    R = rng.standard_normal(r_tot)
    x = rng.standard_normal(p_tot)
    return {"bond_l": bond_l, "bond_r": bond_r, "W": W, "L": L, "R": R, "x": x, "sym": sym}


def c2_problem(seed=2024, total=1024):
    """Inputs of the single H_eff*psi microbenchmark (U(1)xU(1) Hubbard, d=4, w=6).

    Left bond: synthetic_uu_bond(total); right bond: the same shifted by (1,1).
    L, R and psi block entries are N(0,1) from numpy's PCG64 seeded with
    `seed`, drawn in FlatBlockStorage order (L, then R, then psi).
    """
    bl = synthetic_uu_bond(total)
    br = [((a + 1, b + 1), d) for (a, b), d in bl]
    W = hubbard_w(1.0, 4.0)
    _, _, _, l_tot = env_table(SYM_U1U1, bl, W.wl)
    _, _, _, r_tot = env_table(SYM_U1U1, br, W.wr)
    _, _, _, p_tot = block_table(SYM_U1U1, [W.phys, W.phys, bl, br], [DIR_OUT, DIR_OUT, DIR_OUT, DIR_IN])
    rng = np.random.default_rng(seed)
    L = rng.standard_normal(l_tot)
    R = rng.standard_normal(r_tot)
    x = rng.standard_normal(p_tot)
    return {"bond_l": bl, "bond_r": br, "W": W, "L": L, "R": R, "x": x, "sym": SYM_U1U1}


# ---------------------------------------------------------------- chains and states
def chain_mpo(kind, n, **params):
    """Site-by-site MPO of an open chain: [left, bulk, ..., bulk, right]."""
    build = {"heisenberg": heisenberg_w, "hubbard": hubbard_w}[kind]
    return [build(position="left" if i == 0 else "right" if i == n - 1 else "bulk", **params) for i in range(n)]


def _qkey(q):
    return (q, 0) if np.isscalar(q) else tuple(q)


def initial_mps(sym, phys, total, d0, seed=0, min_sector_dim=1):
    """Seeded random MPS with N(0,1) entries in every allowed block.

    Bond i carries the charges reachable from the left with i sites and from
    which `total` is reachable with the remaining sites; sector dimensions are
    proportional to min(left count, right count) of basis configurations,
    capped by that count, at least `min_sector_dim`, rounded to d0 in total by
    largest remainder.  With min_sector_dim=0 the improbable charges (those
    that round to zero) are left out, so the state starts with the
    binomially concentrated charge distribution instead of every reachable
    one.  Entries come from numpy's PCG64 seeded with `seed`, site by site in
    block-table order.  Returns (bonds, data).
    """
    n = len(phys)
    total = _qkey(total)
    sq = [[(_qkey(q), int(d)) for q, d in p] for p in phys]

    def conv(acc, site):
        out = {}
        for q, c in acc.items():
            for s, d in site:
                k = (q[0] + s[0], q[1] + s[1])
                out[k] = out.get(k, 0) + c * d
        return out

    left = [{(0, 0): 1}]
    for i in range(n):
        left.append(conv(left[-1], sq[i]))
    right = [None] * (n + 1)
    right[n] = {(0, 0): 1}
    for i in range(n - 1, -1, -1):
        right[i] = conv(right[i + 1], sq[i])
    bonds = []
    for i in range(n + 1):
        exact = {}
        for q, c in left[i].items():
            rest = (total[0] - q[0], total[1] - q[1])
            if rest in right[i]:
                exact[q] = min(c, right[i][rest])
        qs = sorted(exact)
        if not qs:
            raise ValueError(f"no charge sector reachable on bond {i}")
        cap = [exact[q] for q in qs]
        if sum(cap) <= d0:
            dims = cap
        else:
            tot = float(sum(cap))
            raw = [d0 * c / tot for c in cap]
            dims = [max(min_sector_dim, min(c, int(math.floor(r)))) for c, r in zip(cap, raw)]
            order = sorted(range(len(qs)), key=lambda j: (-(raw[j] - math.floor(raw[j])), qs[j]))
            while sum(dims) < d0:
                moved = False
                for j in order:
                    if sum(dims) >= d0:
                        break
                    if dims[j] < cap[j]:
                        dims[j] += 1
                        moved = True
                if not moved:
                    break
            while sum(dims) > d0:
                j = max(range(len(qs)), key=lambda j: (dims[j], qs[j]))
                if dims[j] <= max(1, min_sector_dim):
                    break
                dims[j] -= 1
        kept = [(q, int(d)) for q, d in zip(qs, dims) if d > 0]
        bonds.append(kept)
    rng = np.random.default_rng(seed)
    data = []
    for i in range(n):
        _, _, _, tot = block_table(sym, [phys[i], bonds[i], bonds[i + 1]], [DIR_OUT, DIR_OUT, DIR_IN])
        data.append(rng.standard_normal(tot))
    return bonds, data


def product_mps(sym, phys, states):
    """Product-state MPS: site i in the first basis state of physical sector
    states[i] (an index into phys[i]).  Every bond has one charge of
    dimension 1.  Returns (bonds, data) like initial_mps."""
    n = len(phys)
    q = (0, 0)
    bonds = [[(q, 1)]]
    for i in range(n):
        s = _qkey(phys[i][states[i]][0])
        q = (q[0] + s[0], q[1] + s[1])
        bonds.append([(q, 1)])
    data = []
    for i in range(n):
        keys, offs, secs, tot = block_table(sym, [phys[i], bonds[i], bonds[i + 1]], [DIR_OUT, DIR_OUT, DIR_IN])
        # single-charge bonds admit exactly the block of the chosen physical charge
        if len(keys) != 1:
            raise ValueError(f"product_mps: site {i} admits {len(keys)} blocks")
        x = np.zeros(tot)
        x[offs[0]] = 1.0
        data.append(x)
    return bonds, data


class Dmrg:
    """tk_dmrg: two-site DMRG with every contraction on the GPU."""

    def __init__(self, ctx, mpo_sites, sym, prec="fp64", schedule=(64,), svd_cutoff=1e-10, davidson_tol=1e-10,
                 max_subspace=60, max_iter=1000, precondition=True, switch_rel=1e-6, degeneracy_tol=1e-12,
                 min_bond_dim=1, eigensolver="davidson", max_krylov=100, low_tail=1000.0):
        self.ctx = ctx
        self.sym = sym
        self.n = len(mpo_sites)
        self.phys0 = list(mpo_sites[0].phys)
        self._keep = []
        sites = []
        for m in mpo_sites:
            p, a, b = QIndexArg(m.phys), QIndexArg(m.wl), QIndexArg(m.wr)
            w = np.ascontiguousarray(m.w, dtype=np.float64)
            self._keep += [p, a, b, w]
            sites.append(tk_mpo_site(p.c, a.c, b.c, w.ctypes.data_as(C.POINTER(C.c_double))))
        arr = (tk_mpo_site * self.n)(*sites)
        sched = np.array(schedule, dtype=np.int64)
        self._keep.append(sched)
        cfg = tk_dmrg_config(PREC[prec], sym, self.n, len(sched), sched.ctypes.data_as(C.POINTER(C.c_int64)),
                             svd_cutoff, min_bond_dim, degeneracy_tol, davidson_tol, max_subspace, max_iter,
                             1 if precondition else 0, switch_rel, {"davidson": 0, "lanczos": 1}[eigensolver],
                             max_krylov, low_tail if low_tail else -1.0)
        h = C.c_void_p()
        _check(ctx.L.tk_dmrg_create(ctx.h, C.byref(cfg), arr, C.byref(h)), ctx.h)
        self.h = h
        # the cancel flag of run_with_cancel_flag (sweep/mod.rs:274-291): the
        # library reads this int32 before every step
        self.cancel_flag = np.zeros(1, dtype=np.int32)
        _check(ctx.L.tk_dmrg_set_cancel_flag(self.h, self.cancel_flag.ctypes.data_as(C.POINTER(C.c_int32))), ctx.h)
        ctx.adopt(self)

    def cancel(self, on=True):
        """Raise (or clear) the cancel flag; a running sweep stops before its
        next step with TK_ERR_CANCELLED.  Safe from another thread."""
        self.cancel_flag[0] = 1 if on else 0

    def subspace_expansion(self, alpha=1e-4, max_new=0, adaptive=True):
        """Subspace expansion of sweep_single_site (expansion_alpha,
        expansion_vectors and adaptive_expansion of TdvpStabilizationConfig,
        ARCHITECTURE.md:1735-1790): each centre move perturbs the two-site
        tensor by alpha times the normalised part of H theta orthogonal to
        theta and splits it with the two-site rule at min(schedule D,
        D + max_new) (max_new <= 0: the schedule's D); adaptive lowers alpha
        tenfold once a sweep's energy change is at its alpha^2 floor.
        alpha = 0 turns it off."""
        _check(self.ctx.L.tk_dmrg_set_subspace_expansion(self.h, float(alpha), int(max_new), 1 if adaptive else 0),
               self.ctx.h)

    def subspace_expansion_alpha(self):
        a = C.c_double(0)
        _check(self.ctx.L.tk_dmrg_subspace_expansion_alpha(self.h, C.byref(a)), self.ctx.h)
        return a.value

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.tk_dmrg_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, bonds, data):
        qs = [QIndexArg(b) for b in bonds]
        arr = (tk_qindex * len(qs))(*[q.c for q in qs])
        ds = [np.ascontiguousarray(d, dtype=np.float64) for d in data]
        ptrs = (C.POINTER(C.c_double) * len(ds))(*[d.ctypes.data_as(C.POINTER(C.c_double)) for d in ds])
        _check(self.ctx.L.tk_dmrg_set_state(self.h, arr, ptrs), self.ctx.h)

    def prepare(self):
        _check(self.ctx.L.tk_dmrg_prepare(self.h), self.ctx.h)

    def sweep(self):
        s = tk_sweep_stats()
        _check(self.ctx.L.tk_dmrg_sweep(self.h, C.byref(s)), self.ctx.h)
        return {f: getattr(s, f) for f, _ in tk_sweep_stats._fields_}

    def sweep_single_site(self):
        """One single-site sweep (dmrg_step_single_site, sweep/mod.rs:451-509, with write-back)."""
        s = tk_sweep_stats()
        _check(self.ctx.L.tk_dmrg_sweep_single_site(self.h, C.byref(s)), self.ctx.h)
        return {f: getattr(s, f) for f, _ in tk_sweep_stats._fields_}

    def apply_site_op(self, op, dq):
        """psi <- O psi for O on site 0 (rows out, columns in, in the caller's order
        of mpo[0].phys) of charge dq; returns |O psi|^2 of the normalised psi."""
        op = np.asarray(op, dtype=np.float64)
        secs = [(_qkey(q), int(dd)) for q, dd in self.phys0]
        order = sorted(range(len(secs)), key=lambda i: secs[i][0])
        off, acc = {}, 0
        for i in order:
            off[i] = acc
            acc += secs[i][1]
        perm = np.array([off[i] + x for i, (_, dd) in enumerate(secs) for x in range(dd)])
        srt = np.zeros_like(op)
        srt[np.ix_(perm, perm)] = op
        srt = np.ascontiguousarray(srt)
        q = np.array(_qkey(dq), dtype=np.int32)
        n2 = C.c_double(0)
        _check(self.ctx.L.tk_dmrg_apply_site_op(self.h, srt.ctypes.data_as(C.POINTER(C.c_double)),
                                                q.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n2)), self.ctx.h)
        return n2.value

    def overlap(self, ket):
        """<self|ket> of two states over the same physical legs (norms not divided out)."""
        v = C.c_double(0)
        _check(self.ctx.L.tk_dmrg_overlap(self.h, ket.h, C.byref(v)), self.ctx.h)
        return v.value

    def greens_function_tau(self, e0, dtau, n_steps, max_bond=0, krylov_dim=30, krylov_tol=1e-12, adaptive_tol=0.0):
        """Imaginary-time G(tau_k) = -<phi| exp(-tau_k (H - e0)) |phi>, tau_k = 2 k dtau,
        k = 0..n_steps, of the current state phi (for example O psi_0 after
        apply_site_op; multiply by its norm^2): each TDVP step of dtau evolves phi to
        exp(-k dtau H) phi, and G(2 k dtau) = -exp(2 sum log_norm + 2 k dtau e0)."""
        taus, g, acc = [0.0], [-1.0], 0.0
        for k in range(1, n_steps + 1):
            acc += self.tdvp_step(dtau, max_bond, krylov_dim, krylov_tol, adaptive_tol)["log_norm"]
            taus.append(2.0 * k * dtau)
            g.append(-math.exp(2.0 * acc + 2.0 * k * dtau * e0))
        return np.array(taus), np.array(g)

    def tdvp_step(self, tau, max_bond=0, krylov_dim=30, krylov_tol=1e-12, adaptive_tol=0.0):
        """One second-order two-site TDVP step psi <- exp(-tau H) psi, renormalised
        (TdvpDriver::step, tdvp/mod.rs:91-141); adaptive_tol > 0 lets each Krylov
        exponential stop early (krylov_dim becomes a maximum)."""
        s = tk_tdvp_stats()
        _check(self.ctx.L.tk_dmrg_tdvp_step(self.h, float(tau), int(max_bond), int(krylov_dim), float(krylov_tol),
                                            float(adaptive_tol), C.byref(s)), self.ctx.h)
        return {f: getattr(s, f) for f, _ in tk_tdvp_stats._fields_ if f != "pad"}

    def steps(self):
        n = C.c_int32(0)
        buf = (tk_step_stats * (2 * self.n + 2))()
        _check(self.ctx.L.tk_dmrg_steps(self.h, len(buf), C.byref(n), buf), self.ctx.h)
        return [{f: getattr(buf[k], f) for f, _ in tk_step_stats._fields_} for k in range(n.value)]

    def bond(self, i):
        n = C.c_int32(0)
        qn = np.zeros(2 * 4096, dtype=np.int32)
        dims = np.zeros(4096, dtype=np.int64)
        _check(self.ctx.L.tk_dmrg_bond(self.h, i, 4096, C.byref(n), qn.ctypes.data_as(C.POINTER(C.c_int32)),
                                       dims.ctypes.data_as(C.POINTER(C.c_int64))), self.ctx.h)
        return [((int(qn[2 * k]), int(qn[2 * k + 1])), int(dims[k])) for k in range(n.value)]

    def site(self, i):
        n = C.c_int64(0)
        self.ctx.L.tk_dmrg_site(self.h, i, 0, C.byref(n), None)
        out = np.zeros(n.value)
        _check(self.ctx.L.tk_dmrg_site(self.h, i, n.value, C.byref(n), out.ctypes.data_as(C.POINTER(C.c_double))),
               self.ctx.h)
        return out

    def checkpoint(self, sweep_index, energy, stats=None):
        """The engine's MPS (every bond QIndex and site tensor) with the
        reference checkpoint's metadata (tk-dmrg/src/checkpoint.rs:16-24)."""
        bonds = [self.bond(i) for i in range(self.n + 1)]
        sites = [self.site(i) for i in range(self.n)]
        return DMRGCheckpoint(int(sweep_index), float(energy), dict(stats or {}), bonds, sites)

    def restore(self, cp, prepare=True):
        """Resume from a DMRGCheckpoint: the saved MPS becomes the state
        (then right-canonicalised and the environments rebuilt unless
        prepare=False)."""
        if len(cp.sites) != self.n or len(cp.bonds) != self.n + 1:
            raise ValueError(f"checkpoint of {len(cp.sites)} sites for a chain of {self.n}")
        self.set_state(cp.bonds, cp.sites)
        if prepare:
            self.prepare()


@dataclass
class DMRGCheckpoint:
    """DMRG engine state on disk.  The reference's DMRGCheckpoint
    (tk-dmrg/src/checkpoint.rs:16-40) holds sweep_index, energy and stats and
    leaves the tensors to a later serde format; here the MPS is stored too:
    every bond as (charges, dims) and every site tensor in FlatBlockStorage
    order.  File format: numpy .npz without pickled objects (arrays
    n, sweep_index, energy, stats_keys, stats_values, bond_<i>_q (k x 2
    int32), bond_<i>_d (int64), site_<i> (float64)), written atomically to a
    temporary file and renamed, as the reference's write_to_file."""
    sweep_index: int
    energy: float
    stats: dict
    bonds: list
    sites: list

    def write_to_file(self, path):
        arrs = {"n": np.array(len(self.sites), dtype=np.int64),
                "sweep_index": np.array(self.sweep_index, dtype=np.int64),
                "energy": np.array(self.energy, dtype=np.float64),
                "stats_keys": np.array(sorted(self.stats), dtype=np.str_),
                "stats_values": np.array([float(self.stats[k]) for k in sorted(self.stats)], dtype=np.float64)}
        for i, b in enumerate(self.bonds):
            arrs[f"bond_{i}_q"] = np.array([list(_qkey(q)) for q, _ in b], dtype=np.int32).reshape(-1, 2)
            arrs[f"bond_{i}_d"] = np.array([int(d) for _, d in b], dtype=np.int64)
        for i, x in enumerate(self.sites):
            arrs[f"site_{i}"] = np.ascontiguousarray(x, dtype=np.float64)
        path = os.fspath(path)
        tmp = path + ".tmp"
        with open(tmp, "wb") as fh:
            np.savez(fh, **arrs)
        os.replace(tmp, path)

    @staticmethod
    def read_from_file(path):
        with np.load(os.fspath(path), allow_pickle=False) as z:
            n = int(z["n"])
            bonds = [[((int(q[0]), int(q[1])), int(d)) for q, d in zip(z[f"bond_{i}_q"], z[f"bond_{i}_d"])]
                     for i in range(n + 1)]
            sites = [np.array(z[f"site_{i}"]) for i in range(n)]
            stats = {str(k): float(v) for k, v in zip(z["stats_keys"], z["stats_values"])}
            return DMRGCheckpoint(int(z["sweep_index"]), float(z["energy"]), stats, bonds, sites)
