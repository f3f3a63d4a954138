"""ctypes binding of the FP64 CPU oracle (oracle/build/libtk_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  See
oracle/tk_oracle.hpp for the reference file each function restates and the
checks that pin it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libtk_oracle.so")


class or_qindex(C.Structure):
    _fields_ = [("n_sectors", C.c_int32), ("qn", C.POINTER(C.c_int32)), ("dims", C.POINTER(C.c_int64))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
    return _lib


class Q:
    """Keeps a qindex's arrays alive; sectors as ((a, b), dim) or (q, dim)."""

    def __init__(self, sectors):
        self.sectors = [((q, 0) if np.isscalar(q) else tuple(q), int(d)) for q, d in sectors]
        self.qn = np.array([c for q, _ in self.sectors for c in q], dtype=np.int32)
        self.dims = np.array([d for _, d in self.sectors], dtype=np.int64)
        self.c = or_qindex(len(self.sectors), self.qn.ctypes.data_as(C.POINTER(C.c_int32)),
                           self.dims.ctypes.data_as(C.POINTER(C.c_int64)))

    def sorted(self):
        return sorted(self.sectors)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def pack_key(sym, qns):
    a = np.array([c for q in qns for c in ((q, 0) if np.isscalar(q) else q)], dtype=np.int32)
    k = C.c_uint64(0)
    st = lib().or_pack_key(sym, len(qns), _p(a, C.c_int32), C.byref(k))
    if st:
        raise ValueError("pack failed")
    return k.value


def unpack_key(sym, key, rank):
    """PackedSectorKey::unpack (sector_key.rs:48-55), pure Python."""
    w = 8 if sym == 1 else 16
    out = []
    for i in range(rank):
        bits = (int(key) >> (i * w)) & ((1 << w) - 1)
        a = bits & 0xFF
        a = a - 256 if a >= 128 else a
        b = (bits >> 8) & 0xFF if sym == 2 else 0
        b = b - 256 if b >= 128 else b
        out.append((a, b))
    return out


def block_table(sym, legs, dirs, flux=(0, 0)):
    qs = [Q(l) for l in legs]
    arr = (or_qindex * len(qs))(*[q.c for q in qs])
    d = np.array(dirs, dtype=np.int32)
    f = np.array(flux, dtype=np.int32)
    n = C.c_int64(0)
    tot = C.c_int64(0)
    lib().or_block_table(sym, len(qs), arr, _p(d, C.c_int32), _p(f, C.c_int32), 0, C.byref(n), None, None,
                         C.byref(tot))
    keys = np.zeros(n.value, dtype=np.uint64)
    offs = np.zeros(n.value, dtype=np.int64)
    st = lib().or_block_table(sym, len(qs), arr, _p(d, C.c_int32), _p(f, C.c_int32), n.value, C.byref(n),
                              _p(keys, C.c_uint64), _p(offs, C.c_int64), C.byref(tot))
    if st:
        raise RuntimeError(f"or_block_table failed {st}")
    return keys, offs, tot.value


def fuse_map(sym, legs, dirs):
    qs = [Q(l) for l in legs]
    arr = (or_qindex * len(qs))(*[q.c for q in qs])
    d = np.array(dirs, dtype=np.int32)
    cap = int(np.prod([len(l) for l in legs]))
    nf = C.c_int32(0)
    nc = C.c_int64(0)
    fq = np.zeros(2 * cap, dtype=np.int32)
    fd = np.zeros(cap, dtype=np.int64)
    cs = np.zeros(cap, dtype=np.int32)
    co = np.zeros(cap, dtype=np.int64)
    st = lib().or_fuse_map(sym, len(qs), arr, _p(d, C.c_int32), cap, C.byref(nf), _p(fq, C.c_int32),
                           _p(fd, C.c_int64), C.byref(nc), _p(cs, C.c_int32), _p(co, C.c_int64))
    if st:
        raise RuntimeError("or_fuse_map failed")
    fused = [((int(fq[2 * i]), int(fq[2 * i + 1])), int(fd[i])) for i in range(nf.value)]
    return fused, cs[:nc.value].copy(), co[:nc.value].copy()


def block_gemm(sym, a_rows, a_cols, a_flux, a_data, b_rows, b_cols, b_flux, b_data):
    qa_r, qa_c, qb_r, qb_c = Q(a_rows), Q(a_cols), Q(b_rows), Q(b_cols)
    fa = np.array(a_flux, dtype=np.int32)
    fb = np.array(b_flux, dtype=np.int32)
    ad = np.ascontiguousarray(a_data, dtype=np.float64)
    bd = np.ascontiguousarray(b_data, dtype=np.float64)
    cap = 4096
    keys = np.zeros(cap, dtype=np.uint64)
    offs = np.zeros(cap, dtype=np.int64)
    cdat = np.zeros(1 << 20)
    n = C.c_int64(0)
    tot = C.c_int64(0)
    st = lib().or_block_gemm(sym, C.byref(qa_r.c), C.byref(qa_c.c), _p(fa, C.c_int32), _p(ad, C.c_double),
                             C.byref(qb_r.c), C.byref(qb_c.c), _p(fb, C.c_int32), _p(bd, C.c_double), cap,
                             C.byref(n), _p(keys, C.c_uint64), _p(offs, C.c_int64), _p(cdat, C.c_double),
                             C.byref(tot))
    if st:
        raise RuntimeError(f"or_block_gemm failed {st}")
    return keys[:n.value].copy(), offs[:n.value].copy(), cdat[:tot.value].copy()


def _heff_args(p):
    W = p["W"]
    keep = [Q(W.phys), Q(W.phys), Q(p["bond_l"]), Q(p["bond_r"]), Q(W.wl), Q(W.wr), Q(W.wr)]
    return keep


def heff_apply(p, x, nthreads=0, Wi=None, Wj=None):
    Wi = Wi if Wi is not None else p["W"]
    Wj = Wj if Wj is not None else p["W"]
    keep = [Q(Wi.phys), Q(Wj.phys), Q(p["bond_l"]), Q(p["bond_r"]), Q(Wi.wl), Q(Wi.wr), Q(Wj.wr)]
    wi = np.ascontiguousarray(Wi.w, dtype=np.float64)
    wj = np.ascontiguousarray(Wj.w, dtype=np.float64)
    L = np.ascontiguousarray(p["L"])
    R = np.ascontiguousarray(p["R"])
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    st = lib().or_heff_apply(p["sym"], *[C.byref(k.c) for k in keep], _p(wi, C.c_double), _p(wj, C.c_double),
                             _p(L, C.c_double), _p(R, C.c_double), _p(x, C.c_double), _p(y, C.c_double),
                             int(nthreads))
    if st:
        raise RuntimeError(f"or_heff_apply failed {st}")
    return y


def heff_diag(p, Wi=None, Wj=None):
    Wi = Wi if Wi is not None else p["W"]
    Wj = Wj if Wj is not None else p["W"]
    keep = [Q(Wi.phys), Q(Wj.phys), Q(p["bond_l"]), Q(p["bond_r"]), Q(Wi.wl), Q(Wi.wr), Q(Wj.wr)]
    wi = np.ascontiguousarray(Wi.w, dtype=np.float64)
    wj = np.ascontiguousarray(Wj.w, dtype=np.float64)
    L = np.ascontiguousarray(p["L"])
    R = np.ascontiguousarray(p["R"])
    n = len(p["x"])
    d = np.zeros(n)
    st = lib().or_heff_diag(p["sym"], *[C.byref(k.c) for k in keep], _p(wi, C.c_double), _p(wj, C.c_double),
                            _p(L, C.c_double), _p(R, C.c_double), _p(d, C.c_double))
    if st:
        raise RuntimeError(f"or_heff_diag failed {st}")
    return d


def heff_dense_reference(p, x_dense, Wi=None, Wj=None):
    Wi = Wi if Wi is not None else p["W"]
    Wj = Wj if Wj is not None else p["W"]
    keep = [Q(Wi.phys), Q(Wj.phys), Q(p["bond_l"]), Q(p["bond_r"]), Q(Wi.wl), Q(Wi.wr), Q(Wj.wr)]
    wi = np.ascontiguousarray(Wi.w, dtype=np.float64)
    wj = np.ascontiguousarray(Wj.w, dtype=np.float64)
    L = np.ascontiguousarray(p["L"])
    R = np.ascontiguousarray(p["R"])
    xd = np.ascontiguousarray(x_dense, dtype=np.float64)
    yd = np.zeros_like(xd)
    st = lib().or_heff_dense_reference(p["sym"], *[C.byref(k.c) for k in keep], _p(wi, C.c_double),
                                       _p(wj, C.c_double), _p(L, C.c_double), _p(R, C.c_double),
                                       _p(xd, C.c_double), _p(yd, C.c_double))
    if st:
        raise RuntimeError(f"or_heff_dense_reference failed {st}")
    return yd


def psi_to_dense(sym, phys_i, phys_j, bond_l, bond_r, flat):
    """Scatter a FlatBlockStorage two-site vector into the dense full-index
    (s_i, s_j, a_L, a_R) array of the reference layout (environments/mod.rs:489-490)."""
    legs = [sorted(Q(l).sectors) for l in (phys_i, phys_j, bond_l, bond_r)]
    offs = []
    for l in legs:
        o, acc = {}, 0
        for q, d in l:
            o[q] = (acc, d)
            acc += d
        offs.append((o, acc))
    keys, boffs, tot = block_table(sym, [phys_i, phys_j, bond_l, bond_r], [1, 1, 1, 0])
    dense = np.zeros(tuple(t for _, t in offs))
    for k, off in zip(keys, boffs):
        qs = unpack_key(sym, k, 4)
        sl = [offs[i][0][qs[i]] for i in range(4)]
        shape = tuple(d for _, d in sl)
        n = int(np.prod(shape))
        dense[tuple(slice(o, o + d) for o, d in sl)] = flat[off:off + n].reshape(shape)
    return dense


def dense_to_psi(sym, phys_i, phys_j, bond_l, bond_r, dense):
    legs = [sorted(Q(l).sectors) for l in (phys_i, phys_j, bond_l, bond_r)]
    offs = []
    for l in legs:
        o, acc = {}, 0
        for q, d in l:
            o[q] = (acc, d)
            acc += d
        offs.append(o)
    keys, boffs, tot = block_table(sym, [phys_i, phys_j, bond_l, bond_r], [1, 1, 1, 0])
    flat = np.zeros(tot)
    for k, off in zip(keys, boffs):
        qs = unpack_key(sym, k, 4)
        sl = [offs[i][qs[i]] for i in range(4)]
        blk = dense[tuple(slice(o, o + d) for o, d in sl)]
        flat[off:off + blk.size] = blk.ravel()
    return flat


def _sorted_leg(leg):
    return sorted(Q(leg).sectors)


def block_permute(sym, legs, dirs, flux, perm, x):
    """BlockSparseTensor::permute (block_sparse.rs:400-431): unpack each key,
    permute the quantum numbers and the block, repack, re-sort (the output
    table is the permuted legs' table)."""
    rank = len(legs)
    keys, offs, tot = block_table(sym, legs, dirs, flux)
    okeys, ooffs, otot = block_table(sym, [legs[i] for i in perm], [dirs[i] for i in perm], flux)
    where = {int(k): int(o) for k, o in zip(okeys, ooffs)}
    dims_of = [dict(_sorted_leg(l)) for l in legs]
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros(otot)
    for k, off in zip(keys, offs):
        qs = unpack_key(sym, k, rank)
        dims = [dims_of[i][qs[i]] for i in range(rank)]
        n = int(np.prod(dims))
        blk = x[off:off + n].reshape(dims)
        o = where[pack_key(sym, [qs[i] for i in perm])]
        out[o:o + n] = blk.transpose(perm).ravel()
    return out


def fused_legs(sym, legs, dirs, start, end):
    """Legs and directions after fuse_legs(start..end): the fused QIndex of the
    fuse map with direction Incoming (block_sparse.rs:522-537); a single leg
    is unchanged (block_sparse.rs:457-468)."""
    if end - start == 1:
        return list(legs), list(dirs)
    fused, _, _ = fuse_map(sym, legs[start:end], dirs[start:end])
    return list(legs[:start]) + [fused] + list(legs[end:]), list(dirs[:start]) + [0] + list(dirs[end:])


def _fuse_moves(sym, legs, dirs, flux, start, end):
    """(original offset, fused offset, outer, chunk, fused stride) per block:
    each original block is `outer` chunks of fuse_dim * trail elements that sit
    at offset * trail inside rows of fused_total * trail (block_sparse.rs:583-600)."""
    rank = len(legs)
    fused, cs, co = fuse_map(sym, legs[start:end], dirs[start:end])
    subs = [_sorted_leg(l) for l in legs[start:end]]
    pos = [{q: i for i, (q, _) in enumerate(sl)} for sl in subs]
    radix = [len(sl) for sl in subs]
    nl, nd = fused_legs(sym, legs, dirs, start, end)
    keys, offs, tot = block_table(sym, legs, dirs, flux)
    okeys, ooffs, otot = block_table(sym, nl, nd, flux)
    where = {int(k): int(o) for k, o in zip(okeys, ooffs)}
    dims_of = [dict(_sorted_leg(l)) for l in legs]
    moves = []
    for k, off in zip(keys, offs):
        qs = unpack_key(sym, k, rank)
        dims = [dims_of[i][qs[i]] for i in range(rank)]
        c = 0
        for l in range(start, end):
            c = c * radix[l - start] + pos[l - start][qs[l]]
        fq, ftot = fused[int(cs[c])]
        o = where[pack_key(sym, list(qs[:start]) + [fq] + list(qs[end:]))]
        outer = int(np.prod(dims[:start])) if start > 0 else 1
        fdim = int(np.prod(dims[start:end]))
        trail = int(np.prod(dims[end:])) if end < rank else 1
        moves.append((int(off), o + int(co[c]) * trail, outer, fdim * trail, ftot * trail))
    return moves, tot, otot


def block_fuse(sym, legs, dirs, flux, start, end, x):
    """BlockSparseTensor::fuse_legs(start..end) (block_sparse.rs:444-611)."""
    x = np.asarray(x, dtype=np.float64)
    if end - start == 1:
        return x.copy()
    moves, tot, otot = _fuse_moves(sym, legs, dirs, flux, start, end)
    out = np.zeros(otot)
    for src, dst, outer, chunk, fstride in moves:
        for o in range(outer):
            out[dst + o * fstride:dst + o * fstride + chunk] = x[src + o * chunk:src + (o + 1) * chunk]
    return out


def block_split(sym, legs, dirs, flux, start, end, xf):
    """BlockSparseTensor::split_leg (block_sparse.rs:622-793): the inverse of
    block_fuse for the same original legs."""
    xf = np.asarray(xf, dtype=np.float64)
    if end - start == 1:
        return xf.copy()
    moves, tot, otot = _fuse_moves(sym, legs, dirs, flux, start, end)
    out = np.zeros(tot)
    for src, dst, outer, chunk, fstride in moves:
        for o in range(outer):
            out[src + o * chunk:src + (o + 1) * chunk] = xf[dst + o * fstride:dst + o * fstride + chunk]
    return out


def to_dense(sym, legs, dirs, flux, x):
    """Full dense array of a FlatBlockStorage tensor: sectors in QIndex order
    along every leg (test helper)."""
    rank = len(legs)
    offs = []
    for l in legs:
        o, acc = {}, 0
        for q, d in _sorted_leg(l):
            o[q] = (acc, d)
            acc += d
        offs.append((o, acc))
    keys, boffs, tot = block_table(sym, legs, dirs, flux)
    dense = np.zeros(tuple(t for _, t in offs))
    for k, off in zip(keys, boffs):
        qs = unpack_key(sym, k, rank)
        sl = [offs[i][0][qs[i]] for i in range(rank)]
        shape = tuple(d for _, d in sl)
        n = int(np.prod(shape))
        dense[tuple(slice(o, o + d) for o, d in sl)] = np.asarray(x[off:off + n]).reshape(shape)
    return dense


def qindex_layout(sectors):
    """QIndex::new + iter_sectors offsets (sector_key.rs:108-158)."""
    q = Q(sectors)
    n = len(q.sectors)
    qn = np.zeros(2 * n, dtype=np.int32)
    dims = np.zeros(n, dtype=np.int64)
    offs = np.zeros(n, dtype=np.int64)
    tot = C.c_int64(0)
    if lib().or_qindex_layout(C.byref(q.c), _p(qn, C.c_int32), _p(dims, C.c_int64), _p(offs, C.c_int64),
                              C.byref(tot)):
        raise RuntimeError("or_qindex_layout failed")
    return [((int(qn[2 * i]), int(qn[2 * i + 1])), int(dims[i])) for i in range(n)], offs, tot.value


def jacobi_small(h):
    """symmetric_eigen_small (eigensolver/mod.rs:532-601): (eigenvalues, eigenvectors as columns)."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    n = h.shape[0]
    ev = np.zeros(n)
    vv = np.zeros(n * n)
    lib().or_jacobi_small(n, _p(h, C.c_double), _p(ev, C.c_double), _p(vv, C.c_double))
    return ev, vv.reshape(n, n)


def davidson_dense(h, diag=None, x0=None, tol=1e-10, max_subspace=60):
    """Davidson on a dense symmetric matrix; x0=None uses the reference start vector."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    n = h.shape[0]
    d = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
    x = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    ev = C.c_double(0)
    out = np.zeros(n)
    nmv = C.c_int32(0)
    lib().or_davidson_dense(n, _p(h, C.c_double), None if d is None else _p(d, C.c_double),
                            None if x is None else _p(x, C.c_double), C.c_double(tol), max_subspace, C.byref(ev),
                            _p(out, C.c_double), C.byref(nmv))
    return ev.value, out, nmv.value


def svd(a):
    """One-sided Jacobi SVD: (U m x r, s, V n x r), s descending, r = min(m, n)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    m, n = a.shape
    r = min(m, n)
    u = np.zeros(m * r)
    s = np.zeros(r)
    v = np.zeros(n * r)
    lib().or_svd(C.c_int64(m), C.c_int64(n), _p(a, C.c_double), _p(u, C.c_double), _p(s, C.c_double),
                 _p(v, C.c_double))
    return u.reshape(m, r), s, v.reshape(n, r)


def truncation_select(sigmas, max_bond, cutoff, min_bond=0, degeneracy_tol=0.0, sigma2_floor=0.0):
    """truncate_svd's selection (truncation/mod.rs:56-98) over the singular
    values of several blocks (a list of descending arrays): (keep per block,
    n_kept, n_discarded, relative truncation error)."""
    counts = np.array([len(s) for s in sigmas], dtype=np.int64)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.float64) for s in sigmas])
                                if len(sigmas) else np.zeros(0))
    keep = np.zeros(max(1, len(sigmas)), dtype=np.int64)
    nk, nd, te = C.c_int64(0), C.c_int64(0), C.c_double(0)
    L = lib()
    L.or_truncation_select.argtypes = [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_int64,
                                       C.c_double, C.c_int64, C.c_double, C.c_double, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    if L.or_truncation_select(len(sigmas), _p(counts, C.c_int64), _p(flat, C.c_double), int(max_bond),
                              float(cutoff), int(min_bond), float(degeneracy_tol), float(sigma2_floor),
                              _p(keep, C.c_int64), C.byref(nk), C.byref(nd), C.byref(te)):
        raise RuntimeError("or_truncation_select failed")
    return keep[:len(sigmas)].tolist(), nk.value, nd.value, te.value


class Dmrg:
    """The oracle's two-site DMRG (CPU, FP64) on the same MPO and initial state."""

    def __init__(self, mpo_sites, sym, schedule=(64,), svd_cutoff=1e-10, davidson_tol=1e-10, max_subspace=60,
                 max_iter=1000, precondition=True, degeneracy_tol=1e-12, nthreads=0):
        L = lib()
        L.or_dmrg_create.restype = C.c_void_p
        L.or_dmrg_destroy.argtypes = [C.c_void_p]
        L.or_dmrg_set_state.argtypes = [C.c_void_p, C.POINTER(or_qindex), C.POINTER(C.POINTER(C.c_double))]
        L.or_dmrg_prepare.argtypes = [C.c_void_p]
        L.or_dmrg_sweep.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.or_dmrg_bond.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int64)]
        self.n = len(mpo_sites)
        self._keep = []
        ph, wl, wr, ws = [], [], [], []
        for m in mpo_sites:
            p, a, b = Q(m.phys), Q(m.wl), Q(m.wr)
            w = np.ascontiguousarray(m.w, dtype=np.float64)
            self._keep += [p, a, b, w]
            ph.append(p.c)
            wl.append(a.c)
            wr.append(b.c)
            ws.append(w.ctypes.data_as(C.POINTER(C.c_double)))
        sched = np.array(schedule, dtype=np.int64)
        self._keep.append(sched)
        L.or_dmrg_create.argtypes = [C.c_int32, C.c_int32, C.POINTER(or_qindex), C.POINTER(or_qindex),
                                     C.POINTER(or_qindex), C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int64),
                                     C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_int32]
        n = self.n
        self.h = L.or_dmrg_create(sym, n, (or_qindex * n)(*ph), (or_qindex * n)(*wl), (or_qindex * n)(*wr),
                                  (C.POINTER(C.c_double) * n)(*ws), _p(sched, C.c_int64), len(sched), svd_cutoff,
                                  degeneracy_tol, davidson_tol, max_subspace, max_iter, 1 if precondition else 0,
                                  int(nthreads))
        if not self.h:
            raise RuntimeError("or_dmrg_create failed")

    def close(self):
        if self.h:
            lib().or_dmrg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, bonds, data):
        qs = [Q(b) for b in bonds]
        ds = [np.ascontiguousarray(d, dtype=np.float64) for d in data]
        ptrs = (C.POINTER(C.c_double) * len(ds))(*[_p(d, C.c_double) for d in ds])
        if lib().or_dmrg_set_state(self.h, (or_qindex * len(qs))(*[q.c for q in qs]), ptrs):
            raise RuntimeError("or_dmrg_set_state failed")

    def prepare(self):
        if lib().or_dmrg_prepare(self.h):
            raise RuntimeError("or_dmrg_prepare failed")

    def sweep(self):
        out = np.zeros(8)
        if lib().or_dmrg_sweep(self.h, _p(out, C.c_double)):
            raise RuntimeError("or_dmrg_sweep failed")
        keys = ["energy", "max_truncation_error", "max_bond_dim", "seconds", "matvecs", "t_matvec", "t_svd", "t_env"]
        return dict(zip(keys, out.tolist()))

    def bond(self, i):
        n = C.c_int32(0)
        qn = np.zeros(8192, dtype=np.int32)
        dims = np.zeros(4096, dtype=np.int64)
        if lib().or_dmrg_bond(self.h, i, 4096, C.byref(n), _p(qn, C.c_int32), _p(dims, C.c_int64)):
            raise RuntimeError("or_dmrg_bond failed")
        return [((int(qn[2 * k]), int(qn[2 * k + 1])), int(dims[k])) for k in range(n.value)]
