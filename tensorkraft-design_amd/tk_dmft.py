"""DMFT self-consistency on the device: the host mirror of tk-dmft.

Mirrors the reference's impurity and loop modules (tk-dmft/src/impurity/
{bath,discretize,hamiltonian}.rs, tk-dmft/src/loop/{mod,mixing}.rs) over the
C ABI of libtk_hip.so:

* BathParameters (bath.rs:18-166): uniform, hybridization_function on the
  device (tk_hybridization), hybridization_distance, linear_mix.
* lanczos_tridiagonalize (discretize.rs:54-274) -> tk_lanczos_tridiagonalize;
  star_to_chain applies the same recursion to the poles of a star bath.
* BroydenState (mixing.rs:40-150) -> tk_broyden_*.
* aim_chain_mpo: the Anderson impurity model in chain geometry (the impurity
  at site 0, bath sites 1..n_bath coupled by nearest-neighbour hopping; the
  reference's OpSum, hamiltonian.rs:92-131, couples every bath site to the
  impurity directly).
* ImpuritySolver: ground state by two-site DMRG on the device, the impurity
  Green's function G(i w_n) as a Krylov continued fraction (tk_krylov_tridiag)
  in the N+-1 sectors, evaluated by tk_hybridization in chain form.
* DMFTLoop (loop/mod.rs:107-419): Bethe lattice Delta = t^2 G on the Matsubara
  axis, batched least-squares bath fit (tk_bath_fit), chain mapping, mixing,
  convergence on the Matsubara-axis hybridization distance; an optional
  learned proposal (Mlp) judged by one impurity solve.
* Mlp: the FP32 network of the learned proposal, trained with the device's
  forward, backward and Adam kernels (tk_mlp_*).
"""
import ctypes as C
import json
import math
import os
from dataclasses import asdict, dataclass, field

import numpy as np

import tk_hip
from tk_hip import _check

STAR, CHAIN = 0, 1
_pD = C.POINTER(C.c_double)
_pF = C.POINTER(C.c_float)


def _dp(a):
    return a.ctypes.data_as(_pD)


def matsubara(beta, n):
    """Fermionic Matsubara frequencies w_n = (2n+1) pi / beta, n < n."""
    return (2.0 * np.arange(n) + 1.0) * math.pi / beta


# ---------------------------------------------------------------- hybridization
def hybridization(ctx, geometry, params, z):
    """Delta_p(z_j) for every row p of params ([eps..., v...]) on the device.

    geometry STAR: sum_k v_k^2 / (z - eps_k) (bath.rs:61-83);
    CHAIN: v_0^2 / (z - eps_0 - v_1^2 / (z - eps_1 - ...)).
    Returns a complex array (n_problems, n_z)."""
    params = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    npb, two_nb = params.shape
    if two_nb % 2:
        raise ValueError("params rows must hold eps and v")
    z = np.asarray(z, dtype=np.complex128).ravel()
    dp = ctx.to_device(params)
    dr = ctx.to_device(np.ascontiguousarray(z.real))
    di = ctx.to_device(np.ascontiguousarray(z.imag))
    do = ctx.empty(2 * npb * z.size)
    try:
        _check(ctx.L.tk_hybridization(ctx.h, geometry, npb, two_nb // 2, C.c_void_p(dp.ptr), z.size,
                                      C.c_void_p(dr.ptr), C.c_void_p(di.ptr), C.c_void_p(do.ptr)), ctx.h)
        out = do.get().reshape(npb, z.size, 2)
    finally:
        for x in (dp, dr, di, do):
            x.free()
    return out[..., 0] + 1j * out[..., 1]


class BathParameters:
    """Star bath (eps_k, v_k) (bath.rs:18-28)."""

    def __init__(self, epsilon, v):
        self.epsilon = np.asarray(epsilon, dtype=np.float64).copy()
        self.v = np.asarray(v, dtype=np.float64).copy()
        if self.epsilon.shape != self.v.shape:
            raise ValueError("epsilon and v must have the same length")

    @property
    def n_bath(self):
        return self.epsilon.size

    @staticmethod
    def uniform(n_bath, bandwidth, v0):
        """Energies linearly spaced in [-bandwidth/2, bandwidth/2] (bath.rs:37-55)."""
        half = bandwidth / 2.0
        if n_bath <= 1:
            eps = np.zeros(n_bath)
        else:
            eps = np.array([-half + (k / (n_bath - 1.0)) * bandwidth for k in range(n_bath)])
        return BathParameters(eps, np.full(n_bath, float(v0)))

    def params(self):
        return np.concatenate([self.epsilon, self.v])

    def hybridization_function(self, ctx, omega, broadening, geometry=STAR):
        """Delta(omega + i broadening) (bath.rs:61-83)."""
        omega = np.asarray(omega, dtype=np.float64)
        return hybridization(ctx, geometry, self.params(), omega + 1j * broadening)[0]

    def hybridization_distance(self, ctx, other, omega, broadening, geometry=STAR):
        """||Delta_self - Delta_other||_inf / ||Delta_other||_inf (bath.rs:92-120)."""
        if len(omega) == 0:
            raise ValueError("omega must not be empty")
        ds = self.hybridization_function(ctx, omega, broadening, geometry)
        do = other.hybridization_function(ctx, omega, broadening, geometry)
        mx = float(np.max(np.abs(ds - do)))
        mr = float(np.max(np.abs(do)))
        return mx if mr < np.finfo(float).eps else mx / mr

    def linear_mix(self, other, alpha):
        """(1 - alpha) self + alpha other (bath.rs:124-166)."""
        if self.n_bath != other.n_bath:
            raise ValueError("bath sizes must match for mixing")
        return BathParameters((1.0 - alpha) * self.epsilon + alpha * other.epsilon,
                              (1.0 - alpha) * self.v + alpha * other.v)


# ---------------------------------------------------------------- discretisation
def lanczos_points(x, w, n_bath):
    """The discretiser's Lanczos recursion on weighted points (tk_lanczos_tridiagonalize)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    eps = np.zeros(n_bath)
    v = np.zeros(n_bath)
    _check(tk_hip.lib().tk_lanczos_tridiagonalize(x.size, _dp(x), _dp(w), n_bath, _dp(eps), _dp(v)))
    return eps, v


def lanczos_tridiagonalize(delta_target, omega, n_bath):
    """discretize.rs:54-274 on a uniform real grid: weights -Im Delta d_omega / pi
    (the magnitude of a real input, as the reference does for real T);
    returns BathParameters(eps_k = alpha_k, v = [sqrt(W), beta_1, ...])."""
    omega = np.asarray(omega, dtype=np.float64)
    d = np.asarray(delta_target)
    if d.size != omega.size:
        raise ValueError("delta_target and omega must have the same length")
    dw = abs(omega[1] - omega[0]) if omega.size > 1 else 1.0
    im = np.abs(d.imag) if np.iscomplexobj(d) else np.abs(d)
    eps, v = lanczos_points(omega, im * dw / math.pi, n_bath)
    return BathParameters(eps, v)


def chain_to_star(ctx, chain):
    """The star bath of a chain: poles = eigenvalues of the bath-site
    tridiagonal (eps'_k on the diagonal, v'_1.. off it), weights v'_0^2 times
    the squared first eigenvector components (device eigensolver)."""
    eps, v = (np.asarray(a, dtype=np.float64) for a in chain)
    n = eps.size
    h = np.diag(eps)
    for k in range(1, n):
        h[k - 1, k] = h[k, k - 1] = v[k]
    w, vec = tk_hip.HipBackend(ctx).eigh(h)
    return BathParameters(w, np.abs(v[0] * vec[0, :]))


def star_to_chain(bath):
    """The chain (eps'_k, v'_k) with the same Delta(z) as the star bath: the
    Lanczos recursion on the poles eps_k with weights v_k^2."""
    return lanczos_points(bath.epsilon, bath.v ** 2, bath.n_bath)


class BroydenState:
    """Good Broyden (mixing.rs:40-150) in the library."""

    def __init__(self, history_depth=5):
        h = C.c_void_p()
        _check(tk_hip.lib().tk_broyden_create(history_depth, C.byref(h)))
        self.h = h
        self.n_updates = 0

    def update(self, x, f, alpha):
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(x)
        n = C.c_int32(0)
        _check(tk_hip.lib().tk_broyden_update(self.h, x.size, _dp(x), _dp(f), alpha, _dp(out), C.byref(n)))
        self.n_updates = n.value
        return out

    def __del__(self):
        try:
            if self.h:
                tk_hip.lib().tk_broyden_destroy(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------- AIM chain MPO
def aim_chain_mpo(eps_imp, U, eps_chain, v_chain):
    """MPO of H = eps_imp n_0 + U n_0up n_0dn + sum_k eps'_k n_k
    + sum_k v'_k sum_s (c+_{k-1,s} c_{k,s} + h.c.), sites 0..n_bath (d = 4,
    U(1)xU(1)); the bond conventions of tk_hip.hubbard_w with site-dependent
    hopping -t = v'_k on bond (k-1, k)."""
    n = len(eps_chain) + 1
    cu, cdu = tk_hip.F_C_UP, tk_hip.F_CDAG_UP
    cd = tk_hip.F_UP_PARITY @ tk_hip.F_C_DN
    cdd = cd.T.copy()
    ntot = np.diag([0.0, 1.0, 1.0, 2.0])
    wq = [((0, 0), 2), ((-1, 0), 1), ((1, 0), 1), ((0, -1), 1), ((0, 1), 1)]
    sites = []
    for i in range(n):
        onsite = (eps_imp * ntot + U * tk_hip.F_NUPNDN) if i == 0 else eps_chain[i - 1] * ntot
        hop = v_chain[i - 1] if i > 0 else 0.0
        ops = {(0, 0): tk_hip.F_ID, (1, 1): tk_hip.F_ID, (0, 1): onsite,
               (0, 2): cdu @ tk_hip.F_PARITY, (0, 3): tk_hip.F_PARITY @ cu,
               (0, 4): cdd @ tk_hip.F_PARITY, (0, 5): tk_hip.F_PARITY @ cd,
               (2, 1): hop * cu, (3, 1): hop * cdu, (4, 1): hop * cd, (5, 1): hop * cdd}
        pos = "single" if n == 1 else "left" if i == 0 else "right" if i == n - 1 else "bulk"
        sites.append(tk_hip._assemble(ops, 4, 6, tk_hip.HUBBARD_PHYS, wq, pos, start=0, done=1))
    return sites


# ---------------------------------------------------------------- sector Hamiltonians
_NUP = np.array([0, 1, 0, 1])
_NDN = np.array([0, 0, 1, 1])


def _two_site_terms(v_chain):
    """(bond, coefficient, O_left, O_right) of the chain hopping, in the MPO's
    local-operator convention."""
    cu, cdu = tk_hip.F_C_UP, tk_hip.F_CDAG_UP
    cd = tk_hip.F_UP_PARITY @ tk_hip.F_C_DN
    cdd = cd.T.copy()
    F = tk_hip.F_PARITY
    out = []
    for k, v in enumerate(v_chain):
        for ol, orr in ((cdu @ F, cu), (F @ cu, cdu), (cdd @ F, cd), (F @ cd, cdd)):
            out.append((k, v, ol, orr))
    return out


def sector_basis(n_sites, n_up, n_dn):
    """Product states (site 0 most significant, 4 states per site) with the
    given particle numbers: sorted full-space indices."""
    idx = np.arange(4 ** n_sites, dtype=np.int64)
    digits = (idx[:, None] // (4 ** np.arange(n_sites - 1, -1, -1))) % 4
    keep = (_NUP[digits].sum(1) == n_up) & (_NDN[digits].sum(1) == n_dn)
    return idx[keep]


def sector_csr(eps_imp, U, eps_chain, v_chain, n_up, n_dn):
    """CSR of the AIM chain Hamiltonian in one (N_up, N_dn) sector; built on
    the host like a block table, from the MPO's local operators."""
    n = len(eps_chain) + 1
    basis = sector_basis(n, n_up, n_dn)
    pos = {int(b): i for i, b in enumerate(basis)}
    pw = 4 ** np.arange(n - 1, -1, -1)
    digits = (basis[:, None] // pw) % 4
    onsite_e = np.concatenate([[eps_imp], eps_chain])
    diag = (onsite_e[None, :] * (_NUP + _NDN)[digits]).sum(1) + U * (digits[:, 0] == 3)
    rows, cols, vals = [np.arange(basis.size)], [np.arange(basis.size)], [diag]
    for k, v, ol, orr in _two_site_terms(v_chain):
        sl, sr = digits[:, k], digits[:, k + 1]
        for a in range(4):
            for b in range(4):
                amp = ol[a, sl] * orr[b, sr] * v
                nz = np.nonzero(amp)[0]
                if nz.size == 0:
                    continue
                new = basis[nz] + (a - sl[nz]) * pw[k] + (b - sr[nz]) * pw[k + 1]
                rows.append(np.array([pos[int(x)] for x in new]))
                cols.append(nz)
                vals.append(amp[nz])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    x = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, x = r[order], c[order], x[order]
    rowptr = np.zeros(basis.size + 1, dtype=np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    return basis, rowptr, c.astype(np.int32), x


# ---------------------------------------------------------------- impurity solver
def mps_to_dense(be, dmrg, n_sites):
    """The MPS as a full-space vector (site 0 most significant, caller basis
    order of the physical states): site tensors expanded per block on the
    host, the chain product A_0 A_1 ... by device GEMMs."""
    sym = tk_hip.SYM_U1U1
    psi = None
    for i in range(n_sites):
        bl, br = dmrg.bond(i), dmrg.bond(i + 1)
        data = dmrg.site(i)
        keys, offs, secs, tot = tk_hip.block_table(sym, [tk_hip.HUBBARD_PHYS, bl, br],
                                                   [tk_hip.DIR_OUT, tk_hip.DIR_OUT, tk_hip.DIR_IN])
        ps, _, _ = tk_hip.qindex_layout(sym, tk_hip.HUBBARD_PHYS)
        ls, loff, ltot = tk_hip.qindex_layout(sym, bl)
        rs, roff, rtot = tk_hip.qindex_layout(sym, br)
        caller = {q: j for j, (q, _) in enumerate(tk_hip.HUBBARD_PHYS)}
        A = np.zeros((ltot, 4, rtot))
        for b in range(len(keys)):
            s, l, r = secs[b]
            dl, dr = ls[l][1], rs[r][1]
            blk = data[offs[b]:offs[b] + dl * dr].reshape(dl, dr)
            A[loff[l]:loff[l] + dl, caller[ps[s][0]], roff[r]:roff[r] + dr] = blk
        A2 = A.reshape(ltot, 4 * rtot)
        if psi is None:
            psi = A2.reshape(4, rtot)
        else:
            m = psi.shape[0]
            psi = be.gemm(1.0, psi, A2, 0.0, np.zeros((m, 4 * rtot))).reshape(m * 4, rtot)
    return psi.reshape(-1)


class ImpuritySolver:
    """Ground state by DMRG, G(i w_n) by Krylov continued fractions."""

    def __init__(self, ctx, n_bath, U, eps_imp=None, bond_dim=256, sweeps=6, krylov=120):
        self.ctx = ctx
        self.be = tk_hip.HipBackend(ctx)
        self.n_bath, self.U = n_bath, U
        self.eps_imp = -U / 2.0 if eps_imp is None else eps_imp
        self.bond_dim, self.sweeps, self.krylov = bond_dim, sweeps, krylov
        self.n_solves = 0

    def _dmrg(self, eps_chain, v_chain):
        """Two-site DMRG ground state at half filling (N_up, N_dn) =
        (ceil(n/2), floor(n/2)): (energy, open engine, mpo, total charge)."""
        n = self.n_bath + 1
        mpo = aim_chain_mpo(self.eps_imp, self.U, eps_chain, v_chain)
        phys = [m.phys for m in mpo]
        total = ((n + 1) // 2, n // 2) if n % 2 else (n // 2, n // 2)
        bonds, data = tk_hip.initial_mps(tk_hip.SYM_U1U1, phys, total, self.bond_dim, seed=0, min_sector_dim=0)
        d = tk_hip.Dmrg(self.ctx, mpo, tk_hip.SYM_U1U1, schedule=[self.bond_dim], svd_cutoff=1e-12,
                        davidson_tol=1e-11)
        d.set_state(bonds, data)
        d.prepare()
        e_prev = None
        for _ in range(self.sweeps):
            e = d.sweep()["energy"]
            if e_prev is not None and abs(e - e_prev) < 1e-12 * max(1.0, abs(e)):
                break
            e_prev = e
        self.n_solves += 1
        return e, d, mpo, total

    def ground_state(self, eps_chain, v_chain):
        n = self.n_bath + 1
        e, d, _, total = self._dmrg(eps_chain, v_chain)
        psi = mps_to_dense(self.be, d, n)
        d.close()
        return e, psi, total

    def green_time(self, eps_chain, v_chain, dt=0.05, t_max=20.0, max_bond=None, krylov_dim=30, krylov_tol=1e-12,
                   adaptive_tol=1e-12):
        """Retarded G_up(t) = -i <{c_0up(t), c+_0up}> of the impurity at
        t_k = k dt, k = 0..round(t_max / dt), from the DMRG ground state by
        real-time two-site TDVP (the G(t) stage of the reference's spectral
        pipeline, tk-dmft/src/spectral/tdvp.rs:55-108): the particle part
        n_p g_p(t) from phi_p = c+ psi0 / |c+ psi0| and the hole part
        -n_h conj(g_h(t)) from phi_h = c psi0 / |c psi0|, with
        g(t) = -i e^{i E0 t} <phi| e^{-i H t} |phi> (tk_hip.greens_function_real_time).
        Everything stays in MPS form on the device.  Returns (t, G(t), info)."""
        import time as _time
        t0 = _time.perf_counter()
        e0, d, mpo, total = self._dmrg(eps_chain, v_chain)
        n = self.n_bath + 1
        bonds = [d.bond(i) for i in range(n + 1)]
        sites = [d.site(i) for i in range(n)]
        d.close()
        t_gs = _time.perf_counter() - t0
        n_steps = int(round(t_max / dt))
        mb = self.bond_dim if max_bond is None else max_bond
        sym = tk_hip.SYM_U1U1
        g = None
        info = {"e0": e0, "t_ground_state": t_gs}
        for name, op, dq in (("particle", tk_hip.F_CDAG_UP, (1, 0)), ("hole", tk_hip.F_C_UP, (-1, 0))):
            t1 = _time.perf_counter()
            x = tk_hip.Dmrg(self.ctx, mpo, sym, schedule=[mb])
            x.set_state(bonds, sites)
            x.prepare()
            w = x.apply_site_op(op, dq)
            pb = [x.bond(i) for i in range(n + 1)]
            ps = [x.site(i) for i in range(n)]
            x.close()
            ts, gp = tk_hip.greens_function_real_time(self.ctx, mpo, sym, pb, ps, e0, dt, n_steps, max_bond=mb,
                                                      krylov_dim=krylov_dim, krylov_tol=krylov_tol,
                                                      svd_cutoff=1e-12, adaptive_tol=adaptive_tol)
            part = w * gp if name == "particle" else -w * np.conj(gp)
            g = part if g is None else g + part
            info[f"weight_{name}"] = w
            info[f"t_{name}"] = _time.perf_counter() - t1
        return ts, g, info

    def _fraction(self, op_site0, psi, eps_chain, v_chain, sector):
        n = self.n_bath + 1
        full = psi.reshape(4, -1)
        phi = (op_site0 @ full).reshape(-1)
        basis, rowptr, col, val = sector_csr(self.eps_imp, self.U, eps_chain, v_chain, *sector)
        v0 = phi[basis]
        ctx = self.ctx
        dr, dc, dv, d0 = (ctx.to_device(rowptr), ctx.to_device(col), ctx.to_device(val), ctx.to_device(v0))
        m = min(self.krylov, basis.size)
        al = np.zeros(m)
        bt = np.zeros(m)
        steps = C.c_int32(0)
        try:
            _check(ctx.L.tk_krylov_tridiag(ctx.h, basis.size, C.c_void_p(dr.ptr), C.c_void_p(dc.ptr),
                                           C.c_void_p(dv.ptr), C.c_void_p(d0.ptr), m, _dp(al), _dp(bt),
                                           C.byref(steps)), ctx.h)
        finally:
            for x in (dr, dc, dv, d0):
                x.free()
        k = steps.value
        return al[:k], bt[:k]

    def green(self, eps_chain, v_chain, wn):
        """G_up(i w_n) of the impurity in the ground state (paramagnetic)."""
        e0, psi, (nu, nd) = self.ground_state(eps_chain, v_chain)
        cdag = tk_hip.F_CDAG_UP
        ap, bp = self._fraction(cdag, psi, eps_chain, v_chain, (nu + 1, nd))
        ah, bh = self._fraction(cdag.T, psi, eps_chain, v_chain, (nu - 1, nd))
        z = 1j * np.asarray(wn)
        prm = np.stack([np.concatenate([ap - e0, bp]), np.concatenate([-(ah - e0), bh])]) \
            if ap.size == ah.size else None
        if prm is not None:
            g = hybridization(self.ctx, CHAIN, prm, z)
            return e0, g[0] + g[1]
        gp = hybridization(self.ctx, CHAIN, np.concatenate([ap - e0, bp]), z)[0]
        gh = hybridization(self.ctx, CHAIN, np.concatenate([-(ah - e0), bh]), z)[0]
        return e0, gp + gh


def green_time_many(baths, n_bath, U, device=0, workers=4, bond_dim=256, progress=None, **green_kw):
    """G(t) (ImpuritySolver.green_time) of several independent impurity
    problems (the reference's independent-solve parallelism,
    ARCHITECTURE.md:2523-2525, on one device): `workers` host threads, each
    with its own context (HIP stream and memory pool), take the problems in
    turn, so that the kernels of different solves run concurrently on the
    GPU.  baths: [(eps_chain, v_chain)]; progress(i, info), when given, is
    called from the worker thread after problem i.  Returns [(t, G(t), info)]
    in the order of `baths`."""
    import threading
    out = [None] * len(baths)
    errors = []
    nxt = [0]
    lock = threading.Lock()

    def work():
        ctx = tk_hip.Context(device)
        try:
            solver = ImpuritySolver(ctx, n_bath, U, bond_dim=bond_dim)
            while True:
                with lock:
                    if errors or nxt[0] >= len(baths):
                        return
                    i = nxt[0]
                    nxt[0] += 1
                out[i] = solver.green_time(*baths[i], **green_kw)
                if progress is not None:
                    progress(i, out[i][2])
        except Exception as e:  # reported by the caller's thread
            with lock:
                errors.append(e)
        finally:
            ctx.close()

    threads = [threading.Thread(target=work) for _ in range(max(1, min(int(workers), len(baths))))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return out


# ---------------------------------------------------------------- bath fit
def bath_fit(ctx, wn, delta, start, max_iter=1000):
    """Least-squares star fit of delta (n_problems, n_z) at i w_n (tk_bath_fit).
    start: (n_problems, 2 n_bath) [eps..., v...]; returns (params, cost)."""
    delta = np.atleast_2d(delta)
    start = np.ascontiguousarray(np.atleast_2d(start), dtype=np.float64)
    npb, nz = delta.shape
    d = np.ascontiguousarray(np.stack([delta.real, delta.imag], -1).reshape(npb, 2 * nz))
    dw, dd, dp = ctx.to_device(np.ascontiguousarray(wn, dtype=np.float64)), ctx.to_device(d), ctx.to_device(start)
    cost = np.zeros(npb)
    try:
        _check(ctx.L.tk_bath_fit(ctx.h, npb, start.shape[1] // 2, nz, C.c_void_p(dw.ptr), C.c_void_p(dd.ptr),
                                 C.c_void_p(dp.ptr), max_iter, _dp(cost)), ctx.h)
        out = dp.get().reshape(npb, -1)
    finally:
        for x in (dw, dd, dp):
            x.free()
    return out, cost


# ---------------------------------------------------------------- MLP
class Mlp:
    """FP32 network dims[0] -> ... -> dims[-1] on the device (tk_mlp_*)."""

    def __init__(self, ctx, dims, seed=0):
        self.ctx = ctx
        self.dims = list(dims)
        d = np.array(dims, dtype=np.int32)
        h = C.c_void_p()
        _check(ctx.L.tk_mlp_create(ctx.h, len(dims) - 1, d.ctypes.data_as(C.POINTER(C.c_int32)), seed, C.byref(h)),
               ctx.h)
        self.h = h
        n = C.c_int64(0)
        _check(ctx.L.tk_mlp_n_params(h, C.byref(n)), ctx.h)
        self.n_params = n.value
        ctx.adopt(self)

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.tk_mlp_destroy(self.h)
        self.h = None

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        _check(self.ctx.L.tk_mlp_get_params(self.h, p.ctypes.data_as(_pF)), self.ctx.h)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.ctx.L.tk_mlp_set_params(self.h, p.ctypes.data_as(_pF)), self.ctx.h)

    def forward(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        dx = self.ctx.to_device(x)
        dy = self.ctx.empty(x.shape[0] * self.dims[-1], np.float32)
        try:
            _check(self.ctx.L.tk_mlp_forward(self.h, x.shape[0], C.c_void_p(dx.ptr), C.c_void_p(dy.ptr)), self.ctx.h)
            return dy.get().reshape(x.shape[0], self.dims[-1])
        finally:
            dx.free()
            dy.free()

    def loss_grad(self, x, t, want_grad=True):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        t = np.ascontiguousarray(np.atleast_2d(t), dtype=np.float32)
        dx, dt = self.ctx.to_device(x), self.ctx.to_device(t)
        loss = C.c_float(0)
        g = np.zeros(self.n_params, dtype=np.float32) if want_grad else None
        try:
            _check(self.ctx.L.tk_mlp_loss_grad(self.h, x.shape[0], C.c_void_p(dx.ptr), C.c_void_p(dt.ptr),
                                               C.byref(loss), g.ctypes.data_as(_pF) if want_grad else None),
                   self.ctx.h)
        finally:
            dx.free()
            dt.free()
        return loss.value, g

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(self.ctx.L.tk_mlp_adam(self.h, lr, beta1, beta2, eps), self.ctx.h)

    def train(self, x, t, epochs=200, lr=1e-3, batch=None, seed=0):
        """Adam on mini-batches (full batch by default); returns the loss history."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        t = np.ascontiguousarray(np.atleast_2d(t), dtype=np.float32)
        n = x.shape[0]
        b = n if batch is None else batch
        rng = np.random.default_rng(seed)
        hist = []
        for _ in range(epochs):
            idx = rng.permutation(n) if b < n else np.arange(n)
            tot = 0.0
            for s in range(0, n, b):
                j = idx[s:s + b]
                loss, _ = self.loss_grad(x[j], t[j], want_grad=False)
                self.adam(lr)
                tot += loss * len(j)
            hist.append(tot / n)
        return hist


def features(delta):
    """Network input: Re and Im of Delta(i w_n), interleaved (FP32)."""
    delta = np.atleast_2d(delta)
    return np.stack([delta.real, delta.imag], -1).reshape(delta.shape[0], -1).astype(np.float32)


# ---------------------------------------------------------------- DMFT loop
@dataclass
class DMFTCheckpoint:
    """Serializable DMFT checkpoint, the reference's DMFTCheckpoint
    (tk-dmft/src/loop/mod.rs:449-488): JSON with the same field names,
    written atomically ({path}.tmp, then rename).  The bath is the chain
    bath of the last completed iteration; this loop works on the Matsubara
    axis, so the spectral fields stay empty unless a caller fills them."""
    iteration: int
    converged: bool
    config_json: str
    bath_epsilon: list
    bath_v: list
    spectral_omega: list = field(default_factory=list)
    spectral_values: list = field(default_factory=list)

    def write_to_file(self, path):
        path = os.fspath(path)
        base, _ = os.path.splitext(path)
        tmp = base + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(asdict(self), fh, indent=2)
        os.replace(tmp, path)

    @staticmethod
    def read_from_file(path):
        with open(os.fspath(path)) as fh:
            d = json.load(fh)
        return DMFTCheckpoint(int(d["iteration"]), bool(d["converged"]), str(d["config_json"]),
                              [float(x) for x in d["bath_epsilon"]], [float(x) for x in d["bath_v"]],
                              [float(x) for x in d.get("spectral_omega", [])],
                              [float(x) for x in d.get("spectral_values", [])])


class DMFTLoop:
    """Bethe-lattice DMFT (loop/mod.rs:107-419) on the Matsubara axis.

    One iteration: impurity solve of the current chain bath (DMRG ground
    state, Krylov G) -> Delta_new = t^2 G (Bethe, t = half bandwidth / 2) ->
    star fit of Delta_new (device Levenberg-Marquardt) -> chain mapping F ->
    mixing of the chain parameters.  Convergence: the fixed-point residual
    ||Delta(F(b)) - Delta(b)||_inf / ||Delta(b)||_inf < tol on the Matsubara
    axis (the reference compares consecutive baths, loop/mod.rs:195-205).
    With `proposal`, a network proposes a chain bath from the first
    Delta_new; one impurity solve judges it, and it replaces the current bath
    when its fixed-point residual is smaller.
    """

    def __init__(self, ctx, n_bath=7, U=4.0, half_bandwidth=2.0, beta=50.0, n_matsubara=256, bond_dim=256,
                 mixing="broyden", alpha=0.5, history=5, tol=1e-4, max_iterations=40, checkpoint_path=None):
        self.ctx = ctx
        self.checkpoint_path = checkpoint_path
        self.config = {"n_bath": n_bath, "U": U, "half_bandwidth": half_bandwidth, "beta": beta,
                       "n_matsubara": n_matsubara, "bond_dim": bond_dim, "mixing": mixing, "alpha": alpha,
                       "history": history, "tol": tol, "max_iterations": max_iterations}
        self.n_bath, self.U = n_bath, U
        self.t = half_bandwidth / 2.0
        self.wn = matsubara(beta, n_matsubara)
        self.solver = ImpuritySolver(ctx, n_bath, U, bond_dim=bond_dim)
        self.mixing, self.alpha, self.history_depth = mixing, alpha, history
        self.tol, self.max_iterations = tol, max_iterations
        self.history = []

    def checkpoint(self, iteration, chain, converged):
        """Write the DMFTCheckpoint of a completed iteration when a
        checkpoint path is set (after every iteration, as loop/mod.rs:211-218
        intends)."""
        if self.checkpoint_path is None:
            return
        DMFTCheckpoint(int(iteration), bool(converged), json.dumps(self.config), [float(x) for x in chain[0]],
                       [float(x) for x in chain[1]]).write_to_file(self.checkpoint_path)

    def resume_from_checkpoint(self, cp):
        """The chain bath to restart from (loop/mod.rs:271-300): checks the
        bath size against this loop's."""
        if len(cp.bath_epsilon) != len(cp.bath_v):
            raise ValueError("checkpoint bath_epsilon and bath_v must have the same length")
        if len(cp.bath_epsilon) != self.n_bath:
            raise ValueError(f"checkpoint bath size ({len(cp.bath_epsilon)}) must match the loop's ({self.n_bath})")
        return (np.array(cp.bath_epsilon, dtype=np.float64), np.array(cp.bath_v, dtype=np.float64))

    def delta_chain(self, chain):
        eps, v = chain
        return hybridization(self.ctx, CHAIN, np.concatenate([eps, v]), 1j * self.wn)[0]

    @staticmethod
    def distance(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), np.finfo(float).eps))

    def fit(self, delta, start_star):
        p, cost = bath_fit(self.ctx, self.wn, delta, start_star)
        return BathParameters(p[0, :self.n_bath], p[0, self.n_bath:]), float(cost[0])

    def step(self, chain, star):
        """One impurity solve: (Delta_new, fitted star, F(chain), residual)."""
        _, g = self.solver.green(chain[0], chain[1], self.wn)
        d_new = self.t ** 2 * g
        star_new, _ = self.fit(d_new, star.params())
        target = star_to_chain(star_new)
        res = self.distance(self.delta_chain(target), self.delta_chain(chain))
        return d_new, star_new, target, res

    def initial_chain(self):
        """Chain of the non-interacting Bethe bath: Delta_0 = t^2 G_0 with
        the semicircular G_0(z) = (z - sqrt(z^2 - 4 t^2)) / (2 t^2)."""
        z = 1j * self.wn
        g0 = (z - np.sqrt(z * z - 4 * self.t ** 2)) / (2 * self.t ** 2)
        star, _ = self.fit(self.t ** 2 * g0, BathParameters.uniform(self.n_bath, 4 * self.t, 0.3).params())
        return star_to_chain(star), star

    def solve(self, chain=None, star=None, proposal=None):
        """Iterate to a fixed-point residual below tol.  Returns (chain,
        number of impurity solves, converged).

        With `proposal`, the bath after the first solve is the network's
        proposal from that solve's Delta_new instead of the mixed update; the
        next impurity solve judges it: when its residual is not below the
        first one, the loop falls back to the mixed update of the first
        bath (the judging solve is then the only extra cost)."""
        if chain is None:
            chain, star = self.initial_chain()
        if star is None:
            star = chain_to_star(self.ctx, chain)
        mixer = BroydenState(self.history_depth) if self.mixing == "broyden" else None
        self.history = []
        start = self.solver.n_solves
        pending = None  # (residual, chain, target, star) of the bath the proposal replaced

        def mixed(ch, tg):
            x, f = np.concatenate(ch), np.concatenate(tg)
            xn = mixer.update(x, f, self.alpha) if mixer is not None else (1 - self.alpha) * x + self.alpha * f
            return (xn[:self.n_bath], xn[self.n_bath:])

        for it in range(self.max_iterations):
            d_new, star_new, target, res = self.step(chain, star)
            rec = {"iteration": it, "residual": res, "delta_new": d_new, "target": np.concatenate(target)}
            self.history.append(rec)
            self.checkpoint(it, chain, res < self.tol)
            if pending is not None:
                r0, c0, t0, s0 = pending
                pending = None
                rec["proposal_taken"] = bool(res < r0)
                if not res < r0:
                    if res < self.tol:
                        return chain, self.solver.n_solves - start, True
                    chain, star = mixed(c0, t0), s0
                    continue
            if res < self.tol:
                return chain, self.solver.n_solves - start, True
            if proposal is not None and it == 0:
                prop = proposal(d_new)
                if prop is not None:
                    pending = (res, chain, target, star_new)
                    chain, star = prop, chain_to_star(self.ctx, prop)
                    continue
            chain, star = mixed(chain, target), star_new
        return chain, self.solver.n_solves - start, False


def train_bath_proposal(ctx, runs, dims=(512, 256, 256, 14), epochs=20000, lr=1e-3, lr_final=1e-5, seed=0):
    """The learned bath proposal: an Mlp from Delta_new(i w_n) (every
    iteration of the given converged DMFTLoop runs) to the chain parameters of
    that run's self-consistent bath, inputs and targets standardised over the
    training set.  Full-batch Adam with the learning rate falling
    geometrically from lr to lr_final over ten stages of epochs / 10
    (lr_final=None keeps lr): the proposal has to land within a few 1e-4 of
    the self-consistent bath to save solves, because the mixing restarts its
    history from it.  runs: [(history, converged chain)].  Returns (network,
    proposal function for DMFTLoop.solve, loss history)."""
    xs, ts = [], []
    for hist, chain in runs:
        tgt = np.concatenate(chain)
        for rec in hist:
            xs.append(features(rec["delta_new"])[0])
            ts.append(tgt)
    x = np.array(xs, dtype=np.float64)
    t = np.array(ts, dtype=np.float64)
    xm, xsd = x.mean(0), x.std(0) + 1e-8
    tm, tsd = t.mean(0), t.std(0) + 1e-8
    net = Mlp(ctx, list(dims), seed=seed)
    xn = ((x - xm) / xsd).astype(np.float32)
    tn = ((t - tm) / tsd).astype(np.float32)
    if lr_final is None:
        loss = net.train(xn, tn, epochs=epochs, lr=lr)
    else:
        loss = []
        for k in range(10):
            rate = lr * (lr_final / lr) ** (k / 9.0)
            loss += net.train(xn, tn, epochs=max(1, epochs // 10), lr=rate)
    nb = dims[-1] // 2

    def proposal(delta_new):
        z = ((features(delta_new)[0].astype(np.float64) - xm) / xsd).astype(np.float32)
        y = net.forward(z)[0].astype(np.float64) * tsd + tm
        return (y[:nb], np.abs(y[nb:]))

    return net, proposal, loss
