"""Krylov exponential and two-site TDVP on the GPU, through the C ABI.

exp_krylov: the reference's own tests (tk-dmrg/src/tdvp/mod.rs:410-484), parity
with the oracle's restatement of exp_krylov_f64 (relative 1e-12, same number of
Arnoldi vectors) on partial Krylov spaces, and scipy.linalg.expm where the
Krylov space is complete.

TDVP: the reference's TdvpDriver::step leaves the MPS unchanged, so the step is
pinned by mathematics instead: with every bond at full dimension the two-site
projector-splitting integrator is exact, so k steps must reproduce
expm(-k tau H) psi_0 of the dense Hamiltonian (state to 1e-9, norm factor to
1e-9, energy to 1e-10); long imaginary-time runs reach the ED ground energy.
"""
import numpy as np
import pytest
import scipy.linalg

import tdvp_oracle
import tk_hip
from ed import heisenberg_ground_energy
from mps_dense import dense_mpo, dense_mps, sorted_phys

pytestmark = pytest.mark.gpu


class _Dev:
    def __init__(self, ptr):
        self.ptr = ptr


@pytest.fixture(scope="module")
def be(ctx):
    return tk_hip.HipBackend(ctx)


def _dense_matvec(be, ctx, h):
    n = h.shape[0]
    dh = ctx.to_device(np.ascontiguousarray(h, dtype=np.float64))

    def mv(x, y):
        be.gemm_strided(n, 1, n, 1.0, dh, n, 1, _Dev(x), 1, 1, 0.0, _Dev(y), 1, 1)
    return mv, dh


def test_exp_krylov_reference_cases(be, ctx):
    print("case exp_krylov reference tests", flush=True)
    mv, keep = _dense_matvec(be, ctx, np.eye(5))
    r, m = be.exp_krylov(mv, 5, np.array([1.0, 0, 0, 0, 0]), 0.5, 10, 1e-12)
    assert abs(r[0] - np.exp(0.5)) < 1e-8 and np.all(np.abs(r[1:]) < 1e-10) and m == 1
    d = np.array([1.0, 2.0, 3.0])
    mv, keep = _dense_matvec(be, ctx, np.diag(d))
    r, m = be.exp_krylov(mv, 3, np.ones(3), -0.1, 10, 1e-12)
    assert np.max(np.abs(r - np.exp(-0.1 * d))) < 1e-14
    a = 2 * np.eye(4) - np.eye(4, k=1) - np.eye(4, k=-1)
    mv, keep = _dense_matvec(be, ctx, a)
    v = np.array([1.0, 2, 3, 4])
    r, m = be.exp_krylov(mv, 4, v, 0.0, 10, 1e-12)
    assert np.max(np.abs(r - v)) < 1e-10


# (64, 30, -1.0): |alpha A| ~ 11 takes the scaling-and-squaring branch, whose
# six squarings amplify rounding, hence the looser 1e-11 there
@pytest.mark.parametrize("n,kdim,alpha,rel", [(40, 12, -0.3, 1e-12), (40, 20, 0.15, 1e-12), (64, 30, -1.0, 1e-11),
                                              (40, 40, -0.5, 1e-12)])
def test_exp_krylov_matches_oracle(be, ctx, n, kdim, alpha, rel):
    print(f"case exp_krylov n={n} kdim={kdim} alpha={alpha}", flush=True)
    rng = np.random.default_rng(n + kdim)
    g = rng.standard_normal((n, n))
    a = (g + g.T) / 2
    v = rng.standard_normal(n)
    mv, keep = _dense_matvec(be, ctx, a)
    r, m = be.exp_krylov(mv, n, v, alpha, kdim, 1e-13)
    ro, mo = tdvp_oracle.exp_krylov(lambda x: a @ x, v, alpha, kdim, 1e-13)
    assert m == mo
    assert np.linalg.norm(r - ro) <= rel * np.linalg.norm(ro)
    if kdim == n:
        ref = scipy.linalg.expm(alpha * a) @ v
        assert np.linalg.norm(r - ref) <= 1e-11 * np.linalg.norm(ref)


def test_exp_krylov_zero_vector_and_limits(be, ctx):
    mv, keep = _dense_matvec(be, ctx, np.eye(6))
    r, m = be.exp_krylov(mv, 6, np.zeros(6), -1.0, 6, 1e-12)
    assert m == 0 and not np.any(r)
    with pytest.raises(tk_hip.TkError, match="48"):
        be.exp_krylov(mv, 6, np.ones(6), -1.0, 49, 1e-12)


def _tdvp_case(ctx, kind, n, total, tau, steps, **mp):
    mpo = tk_hip.chain_mpo(kind, n, **mp)
    sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, 256, seed=7)
    psi0 = dense_mps(sym, mpo, bonds, data)
    psi0 /= np.linalg.norm(psi0)
    h, _ = dense_mpo(mpo)
    d = tk_hip.Dmrg(ctx, mpo, sym, schedule=(256,), svd_cutoff=1e-14)
    d.set_state(bonds, data)
    d.prepare()
    log_norm, st = 0.0, None
    for _ in range(steps):
        st = d.tdvp_step(tau, max_bond=256, krylov_dim=40, krylov_tol=1e-14)
        log_norm += st["log_norm"]
    psi = dense_mps(sym, mpo, [d.bond(i) for i in range(n + 1)], [d.site(i) for i in range(n)])
    exact = scipy.linalg.expm(-steps * tau * h) @ psi0
    return psi, exact, log_norm, st, h


@pytest.mark.parametrize("kind,n,total,mp", [("heisenberg", 6, (0, 0), {}), ("heisenberg", 7, (1, 0), {}),
                                              ("hubbard", 4, (2, 2), {"t": 1.0, "U": 4.0}),
                                              ("hubbard", 4, (3, 1), {"t": 1.0, "U": 2.0})])
def test_gpu_tdvp_full_bonds_is_exact(ctx, kind, n, total, mp):
    print(f"case tdvp exact {kind} n={n} total={total}", flush=True)
    tau, steps = 0.1, 4
    psi, exact, log_norm, st, h = _tdvp_case(ctx, kind, n, total, tau, steps, **mp)
    nrm = np.linalg.norm(exact)
    assert abs(log_norm - np.log(nrm)) < 1e-9
    target = exact / nrm
    assert abs(np.linalg.norm(psi) - 1) < 1e-10
    sign = np.sign(psi @ target)
    assert np.max(np.abs(sign * psi - target)) < 1e-9
    assert abs(st["energy"] - target @ h @ target) < 1e-10
    assert st["max_truncation_error"] < 1e-20


def test_gpu_tdvp_imaginary_time_reaches_ground_state(ctx):
    print("case tdvp imaginary time heisenberg n=10", flush=True)
    n = 10
    mpo = tk_hip.chain_mpo("heisenberg", n)
    bonds, data = tk_hip.initial_mps(tk_hip.SYM_U1, [m.phys for m in mpo], (0, 0), 64, seed=3)
    d = tk_hip.Dmrg(ctx, mpo, tk_hip.SYM_U1, schedule=(64,), svd_cutoff=1e-14)
    d.set_state(bonds, data)
    d.prepare()
    # total time 45: the excited weight decays like exp(-gap * t) with the
    # Sz = 0 triplet gap ~ 0.34 of the ten-site chain
    e = [d.tdvp_step(0.75, krylov_dim=40)["energy"] for _ in range(60)]
    assert all(e[k + 1] <= e[k] + 1e-12 for k in range(len(e) - 1))  # imaginary time lowers the energy
    assert abs(e[-1] - heisenberg_ground_energy(n)) < 1e-9


def test_gpu_tdvp_after_two_site_sweeps(ctx):
    """A two-site sweep leaves the centre at site 1; the TDVP step first moves
    it to site 0, so imaginary time from a converged state stays there."""
    print("case tdvp after dmrg hubbard n=6", flush=True)
    from ed import hubbard_ground_energy
    n = 6
    mpo = tk_hip.chain_mpo("hubbard", n, t=1.0, U=4.0)
    bonds, data = tk_hip.initial_mps(tk_hip.SYM_U1U1, [m.phys for m in mpo], (3, 3), 64, seed=2)
    d = tk_hip.Dmrg(ctx, mpo, tk_hip.SYM_U1U1, schedule=(64,), davidson_tol=1e-12, svd_cutoff=1e-14)
    d.set_state(bonds, data)
    d.prepare()
    for _ in range(4):
        d.sweep()
    st = d.tdvp_step(0.2, krylov_dim=30)
    e0 = hubbard_ground_energy(n, 3, 3, 1.0, 4.0)
    assert abs(st["energy"] - e0) < 1e-9
    assert abs(st["log_norm"] + 0.2 * e0) < 1e-8  # exp(-tau H) on an eigenstate scales it by exp(-tau E0)


@pytest.mark.parametrize("kind,n,total,opname,dq,mp", [
    ("hubbard", 4, (2, 2), "Cdagup", (1, 0), {"t": 1.0, "U": 4.0}),
    ("heisenberg", 6, (0, 0), "S+", (2, 0), {}),
])
def test_gpu_greens_function_tau_matches_ed(ctx, kind, n, total, opname, dq, mp):
    """G(tau) = -<psi0| O+ exp(-tau (H - E0)) O |psi0> from the DMRG ground state,
    O applied on site 0 and imaginary-time TDVP (exact at full bond dimension),
    against the same correlator of the densified MPO's ground state."""
    print(f"case greens tau {kind} n={n} op={opname}", flush=True)
    mpo = tk_hip.chain_mpo(kind, n, **mp)
    sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, 256, seed=4)
    d = tk_hip.Dmrg(ctx, mpo, sym, schedule=(256,), davidson_tol=1e-12, svd_cutoff=1e-14)
    d.set_state(bonds, data)
    d.prepare()
    for _ in range(4):
        e0 = d.sweep()["energy"]
    site_type = tk_hip.SITE_TYPES["electron" if kind == "hubbard" else "spin-1/2"]()
    op = site_type.ops[opname]
    norm2 = d.apply_site_op(op, dq)
    taus, g = d.greens_function_tau(e0, 0.05, 6, max_bond=256, krylov_dim=40, krylov_tol=1e-14)
    # dense reference in the sorted basis of the densified MPO
    h, ch = dense_mpo(mpo)
    idx = [k for k, q in enumerate(ch) if q == tuple(total)]
    w, v = np.linalg.eigh(h[np.ix_(idx, idx)])
    assert abs(w[0] - e0) < 1e-10 and w[1] - w[0] > 1e-3
    psi0 = np.zeros(h.shape[0])
    psi0[idx] = v[:, 0]
    perm, _ = sorted_phys(mpo[0].phys)
    ops = np.zeros_like(op)
    ops[np.ix_(perm, perm)] = op
    phi = np.kron(ops, np.eye(h.shape[0] // ops.shape[0])) @ psi0
    assert abs(norm2 - phi @ phi) < 1e-10
    ref = np.array([-(phi @ scipy.linalg.expm(-t * (h - w[0] * np.eye(h.shape[0]))) @ phi) for t in taus])
    assert np.max(np.abs(norm2 * g - ref)) < 1e-9 * np.max(np.abs(ref))


@pytest.mark.parametrize("kind,n,total,mp", [("heisenberg", 6, (0, 0), {}), ("hubbard", 4, (2, 2), {"t": 1.0, "U": 4.0})])
def test_gpu_mps_overlap_matches_dense(ctx, kind, n, total, mp):
    print(f"case overlap {kind} n={n}", flush=True)
    mpo = tk_hip.chain_mpo(kind, n, **mp)
    sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
    phys = [m.phys for m in mpo]
    a = tk_hip.initial_mps(sym, phys, total, 5, seed=11)
    b = tk_hip.initial_mps(sym, phys, total, 12, seed=12)
    da, db = tk_hip.Dmrg(ctx, mpo, sym), tk_hip.Dmrg(ctx, mpo, sym)
    da.set_state(*a)
    db.set_state(*b)
    pa, pb = dense_mps(sym, mpo, *a), dense_mps(sym, mpo, *b)
    assert abs(da.overlap(db) - pa @ pb) <= 1e-12 * np.linalg.norm(pa) * np.linalg.norm(pb)
    assert abs(db.overlap(db) - pb @ pb) <= 1e-12 * (pb @ pb)
    db.prepare()  # canonical gauge, norm 1: the overlap is gauge invariant
    assert abs(da.overlap(db) - pa @ pb / np.linalg.norm(pb)) <= 1e-12 * np.linalg.norm(pa)


def _complex_dense(sym, cmpo, bonds, sites):
    phi = dense_mps(sym, cmpo, bonds, sites)
    half = phi.shape[0] // 2  # the ancilla is the most significant site: |R> block, then |I>
    return phi[:half] + 1j * phi[half:]


def test_gpu_real_time_tdvp_full_bonds_is_exact(ctx):
    """exp(-i H t) psi through the ancilla form: every bond (the ancilla's
    included) at full dimension, so the integrator is exact."""
    print("case real-time tdvp hubbard n=4", flush=True)
    mpo = tk_hip.chain_mpo("hubbard", 4, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    cmpo = tk_hip.complex_time_mpo(mpo)
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in cmpo], (2, 2), 256, seed=9)
    psi0 = _complex_dense(sym, cmpo, bonds, data)
    psi0 /= np.linalg.norm(psi0)
    d = tk_hip.Dmrg(ctx, cmpo, sym, schedule=(256,), svd_cutoff=1e-14)
    d.set_state(bonds, data)
    d.prepare()
    dt, steps, acc = 0.1, 4, 0.0
    for _ in range(steps):
        acc += d.tdvp_step(-dt, 256, 40, 1e-14)["log_norm"]
    assert abs(acc) < 1e-10  # unitary
    psi = _complex_dense(sym, cmpo, [d.bond(i) for i in range(len(cmpo) + 1)], [d.site(i) for i in range(len(cmpo))])
    h, _ = dense_mpo(mpo)
    exact = scipy.linalg.expm(-1j * steps * dt * h) @ psi0
    assert np.max(np.abs(psi - exact)) < 1e-9


def test_gpu_greens_function_real_time_matches_ed(ctx):
    """G(t) = -i e^{i E0 t} <phi|e^{-i H t}|phi>, phi = c+_up psi0 / |.|, from the
    DMRG ground state of a 4-site Hubbard chain.  The start state gets a
    component eps = 1e-6 of a random full-rank state so that every bond is at
    full dimension (the integrator is then exact); that perturbs G by at most
    about 2 eps."""
    print("case greens real time hubbard n=4", flush=True)
    from mps_dense import mps_sum
    mpo = tk_hip.chain_mpo("hubbard", 4, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (2, 2), 256, seed=4)
    d = tk_hip.Dmrg(ctx, mpo, sym, schedule=(256,), davidson_tol=1e-12, svd_cutoff=1e-14)
    d.set_state(bonds, data)
    d.prepare()
    for _ in range(4):
        e0 = d.sweep()["energy"]
    op = tk_hip.SITE_TYPES["electron"]().ops["Cdagup"]
    d.apply_site_op(op, (1, 0))
    pb = [d.bond(i) for i in range(5)]
    ps = [d.site(i) for i in range(4)]
    cmpo = tk_hip.complex_time_mpo(mpo)
    cphys = [m.phys for m in cmpo]
    chi = tk_hip.initial_mps(sym, cphys, (3, 2), 256, seed=21)
    eps = 1e-6 / np.linalg.norm(dense_mps(sym, cmpo, *chi))
    start = mps_sum(sym, cphys, tk_hip.complex_state(pb, ps, "R"), chi, cb=eps)
    t, g = tk_hip.greens_function_real_time(ctx, mpo, sym, pb, ps, e0, 0.1, 10, max_bond=256, krylov_dim=40,
                                            krylov_tol=1e-14, svd_cutoff=1e-14, start=start)
    h, ch = dense_mpo(mpo)
    idx = [k for k, q in enumerate(ch) if q == (2, 2)]
    w, v = np.linalg.eigh(h[np.ix_(idx, idx)])
    psi0 = np.zeros(h.shape[0])
    psi0[idx] = v[:, 0]
    perm, _ = sorted_phys(mpo[0].phys)
    ops = np.zeros_like(op)
    ops[np.ix_(perm, perm)] = op
    phi = np.kron(ops, np.eye(h.shape[0] // 4)) @ psi0
    phi /= np.linalg.norm(phi)
    ref = np.array([-1j * np.exp(1j * w[0] * tt) * (phi @ scipy.linalg.expm(-1j * tt * h) @ phi) for tt in t])
    assert abs(g[0] + 1j) < 3e-6
    assert np.max(np.abs(g - ref)) < 3e-6


def test_exp_krylov_adaptive_stop(be, ctx):
    """The early stop (not in the reference): fewer Arnoldi vectors, same result."""
    print("case exp_krylov adaptive", flush=True)
    rng = np.random.default_rng(64)
    g = rng.standard_normal((64, 64))
    a = (g + g.T) / 2
    v = rng.standard_normal(64)
    mv, keep = _dense_matvec(be, ctx, a)
    ref = scipy.linalg.expm(-0.05 * a) @ v
    r, m = be.exp_krylov(mv, 64, v, -0.05, 40, 1e-14, adaptive_tol=1e-13)
    assert m < 40
    assert np.linalg.norm(r - ref) <= 1e-11 * np.linalg.norm(ref)
    r0, m0 = be.exp_krylov(mv, 64, v, -0.05, 40, 1e-14)
    assert m0 == 40 and np.linalg.norm(r0 - ref) <= 1e-12 * np.linalg.norm(ref)


def test_gpu_tdvp_adaptive_krylov_is_exact(ctx):
    print("case tdvp adaptive krylov heisenberg n=6", flush=True)
    mpo = tk_hip.chain_mpo("heisenberg", 6)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), 256, seed=7)
    psi0 = dense_mps(sym, mpo, bonds, data)
    psi0 /= np.linalg.norm(psi0)
    h, _ = dense_mpo(mpo)
    out = {}
    for tol in (0.0, 1e-13):
        d = tk_hip.Dmrg(ctx, mpo, sym, schedule=(256,), svd_cutoff=1e-14)
        d.set_state(bonds, data)
        d.prepare()
        mv = 0
        for _ in range(4):
            mv += d.tdvp_step(0.1, 256, 40, 1e-14, adaptive_tol=tol)["matvecs"]
        psi = dense_mps(sym, mpo, [d.bond(i) for i in range(7)], [d.site(i) for i in range(6)])
        out[tol] = (psi, mv)
    target = scipy.linalg.expm(-0.4 * h) @ psi0
    target /= np.linalg.norm(target)
    for tol, (psi, mv) in out.items():
        assert np.max(np.abs(np.sign(psi @ target) * psi - target)) < 1e-9
    assert out[1e-13][1] < out[0.0][1]


def test_gpu_tdvp_error_paths(ctx):
    """Misuse is reported through the C ABI's status, not by a crash."""
    print("case tdvp error paths", flush=True)
    mpo = tk_hip.chain_mpo("heisenberg", 4)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), 8, seed=1)
    d = tk_hip.Dmrg(ctx, mpo, sym)
    d.set_state(bonds, data)
    with pytest.raises(tk_hip.TkError, match="prepare"):
        d.tdvp_step(0.1)
    d.prepare()
    with pytest.raises(tk_hip.TkError, match="charge"):
        d.apply_site_op(tk_hip.SITE_TYPES["spin-1/2"]().ops["S+"], (0, 0))  # S+ raises 2Sz by 2, not 0
    with pytest.raises(tk_hip.TkError, match="krylov_dim"):
        d.tdvp_step(0.1, krylov_dim=0)
    other = tk_hip.Dmrg(ctx, tk_hip.chain_mpo("heisenberg", 6), sym)
    b6, d6 = tk_hip.initial_mps(sym, [m.phys for m in tk_hip.chain_mpo("heisenberg", 6)], (0, 0), 8, seed=1)
    other.set_state(b6, d6)
    with pytest.raises(tk_hip.TkError, match="length"):
        d.overlap(other)
    st = d.tdvp_step(0.1)  # the engine still works after the rejected calls
    assert np.isfinite(st["energy"])
