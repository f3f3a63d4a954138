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
