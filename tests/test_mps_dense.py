"""The dense test helpers agree with exact diagonalisation and charge conservation."""
import numpy as np

import tk_hip
from ed import heisenberg_ground_energy, hubbard_ground_energy
from mps_dense import dense_mpo, dense_mps


def test_dense_heisenberg_mpo_ground_energy():
    h, ch = dense_mpo(tk_hip.chain_mpo("heisenberg", 6))
    assert np.abs(h - h.T).max() < 1e-14
    assert abs(np.linalg.eigvalsh(h)[0] - heisenberg_ground_energy(6)) < 1e-12


def test_dense_hubbard_mpo_sector_energy():
    h, ch = dense_mpo(tk_hip.chain_mpo("hubbard", 4, t=1.0, U=4.0))
    idx = [k for k, q in enumerate(ch) if q == (2, 2)]
    other = [k for k, q in enumerate(ch) if q != (2, 2)]
    assert np.abs(h[np.ix_(idx, other)]).max() == 0  # H conserves (N_up, N_dn)
    e = np.linalg.eigvalsh(h[np.ix_(idx, idx)])[0]
    assert abs(e - hubbard_ground_energy(4, 2, 2, 1.0, 4.0)) < 1e-12


def test_dense_mps_lives_in_its_charge_sector():
    # totals that an up/down mix-up of the basis order would move to another sector
    cases = [("heisenberg", 6, (0, 0), {}), ("heisenberg", 6, (2, 0), {}),
             ("hubbard", 4, (2, 2), {"t": 1.0, "U": 4.0}), ("hubbard", 4, (3, 1), {"t": 1.0, "U": 4.0})]
    for kind, n, total, mp in cases:
        mpo = tk_hip.chain_mpo(kind, n, **mp)
        sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
        bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, 64, seed=1)
        psi = dense_mps(sym, mpo, bonds, data)
        _, ch = dense_mpo(mpo)
        outside = [k for k, q in enumerate(ch) if q != tuple(total)]
        assert np.linalg.norm(psi) > 0 and np.abs(psi[outside]).max() == 0


def test_mps_sum_and_from_dense_round_trip():
    from mps_dense import from_dense, mps_sum
    import oracle
    mpo = tk_hip.chain_mpo("hubbard", 4, t=1.0, U=4.0)
    phys = [m.phys for m in mpo]
    sym = tk_hip.SYM_U1U1
    a = tk_hip.initial_mps(sym, phys, (2, 2), 6, seed=1)
    b = tk_hip.initial_mps(sym, phys, (2, 2), 9, seed=2)
    s = mps_sum(sym, phys, a, b, cb=-0.3)
    pa, pb, ps = (dense_mps(sym, mpo, *x) for x in (a, b, s))
    assert np.abs(ps - (pa - 0.3 * pb)).max() < 1e-12
    legs = [phys[1], a[0][1], a[0][2]]
    d = oracle.to_dense(sym, legs, [1, 1, 0], (0, 0), a[1][1])
    assert np.array_equal(from_dense(sym, legs, [1, 1, 0], (0, 0), d), np.asarray(a[1][1]))


def test_complex_time_mpo_is_j_kron_h():
    """The ancilla MPO equals J (x) H with J|R> = -|I>, J|I> = |R>, so that
    exp(t J (x) H) carries exp(-i H t) on psi_R + i psi_I."""
    mpo = tk_hip.chain_mpo("hubbard", 3, t=1.0, U=4.0)
    h, _ = dense_mpo(mpo)
    hc, ch = dense_mpo(tk_hip.complex_time_mpo(mpo))
    J = np.array([[0.0, 1.0], [-1.0, 0.0]])  # rows out (R, I), columns in (R, I)
    assert np.abs(hc - np.kron(J, h)).max() < 1e-14
    # d/dt (psi_R, psi_I) = (H psi_I, -H psi_R): the ancilla form of exp(-i H t)
    rng = np.random.default_rng(0)
    pr, pi = rng.standard_normal(h.shape[0]), rng.standard_normal(h.shape[0])
    flow = hc @ np.concatenate([pr, pi])
    assert np.allclose(flow[:h.shape[0]], h @ pi) and np.allclose(flow[h.shape[0]:], -(h @ pr))


def test_complex_state_layout():
    mpo = tk_hip.chain_mpo("heisenberg", 4)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), 8, seed=2)
    phi = dense_mps(sym, mpo, bonds, data)
    cmpo = tk_hip.complex_time_mpo(mpo)
    for part, half in (("R", 0), ("I", 1)):
        b, s = tk_hip.complex_state(bonds, data, part)
        full = dense_mps(sym, cmpo, b, s)
        n = phi.shape[0]
        assert np.array_equal(full[half * n:(half + 1) * n], phi) and not np.any(full[(1 - half) * n:(2 - half) * n])
