"""OpSum and its MPO compiler (CPU only).

* the compiled Heisenberg and Hubbard chains equal the hand-built MPOs
  (densified), which ED and the U=0 closed form pin;
* the compiled cylinder equals the bond-list MPO;
* a Hubbard ladder with hops across the snake (Jordan-Wigner strings over
  other sites): oracle DMRG vs exact diagonalisation;
* the reference's OpSum bookkeeping (push_term, n_terms, scale, extend) and
  the rejected forms.
"""
import numpy as np
import pytest

import oracle
import tk_hip
from ed import free_fermion_energy, hubbard_bonds_ground_energy, hubbard_ground_energy
from test_cylinder_mpo import _dense_from_mpo


def test_opsum_bookkeeping():
    os_ = tk_hip.OpSum(4)
    os_.push_term(1.0, ("Sz", 0), ("Sz", 1)).push_term(0.5, ("S+", 1), ("S-", 2))
    assert os_.n_terms() == 2
    os_.scale(2.0)
    assert [c for c, _ in os_.iter_terms()] == [2.0, 1.0]
    other = tk_hip.OpSum(4)
    other.push_term(0.3, ("Sz", 3))
    os_.extend(other)
    assert os_.n_terms() == 3
    with pytest.raises(ValueError):
        os_.push_term(1.0, ("Sz", 4))  # outside the lattice (opsum.rs:76-86)


def test_compiled_heisenberg_chain_equals_hand_built():
    n = 6
    m = tk_hip.compile_mpo(tk_hip.heisenberg_opsum([(i, i + 1) for i in range(n - 1)], n), n)
    assert all(w.w.shape[3] == 5 for w in m[:-1])
    assert np.abs(_dense_from_mpo(m) - _dense_from_mpo(tk_hip.chain_mpo("heisenberg", n))).max() < 1e-14


def test_compiled_hubbard_chain_equals_hand_built():
    n = 4
    m = tk_hip.compile_mpo(tk_hip.hubbard_opsum([(i, i + 1) for i in range(n - 1)], n, 1.0, 4.0), n, "electron")
    ref = tk_hip.chain_mpo("hubbard", n, t=1.0, U=4.0)
    assert np.abs(_dense_from_mpo(m) - _dense_from_mpo(ref)).max() < 1e-13


def test_compiled_cylinder_equals_bond_mpo():
    b = tk_hip.cylinder_bonds(4, 2)
    m = tk_hip.compile_mpo(tk_hip.heisenberg_opsum(b, 8), 8)
    assert np.abs(_dense_from_mpo(m) - _dense_from_mpo(tk_hip.bond_heisenberg_mpo(8, b))).max() < 1e-14


def test_rejected_terms():
    with pytest.raises(ValueError):  # one fermion operator: parity odd
        tk_hip.compile_mpo(tk_hip.OpSum(3).push_term(1.0, ("Cup", 0), ("Nup", 2)), 3, "electron")
    with pytest.raises(ValueError):
        tk_hip.compile_mpo(tk_hip.OpSum(3).push_term(1.0, ("Sz", 0), ("Sz", 1), ("Sz", 2)), 3)
    with pytest.raises(ValueError):
        tk_hip.compile_mpo(tk_hip.OpSum(3).push_term(1.0, ("Q", 0)), 3)


def _dmrg(mpo, sym, total, sched, sweeps):
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, 16, 0)
    d = oracle.Dmrg(mpo, sym, schedule=sched, davidson_tol=1e-12, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    return [d.sweep() for _ in range(sweeps)]


def test_compiled_hubbard_ladder_matches_ed():
    # 3 x 2 ladder in snake order: the rungs between rows cross other sites
    n = 6
    bonds = tk_hip.cylinder_bonds(2, 3)
    assert max(j - i for i, j in bonds) == 3
    mpo = tk_hip.compile_mpo(tk_hip.hubbard_opsum(bonds, n, 1.0, 4.0), n, "electron")
    st = _dmrg(mpo, tk_hip.SYM_U1U1, (3, 3), (64, 128), 5)
    assert abs(st[-1]["energy"] - hubbard_bonds_ground_energy(n, bonds, 3, 3, 1.0, 4.0)) < 1e-9


def test_ed_bonds_matches_chain_ed():
    n = 6
    chain = [(i, i + 1) for i in range(n - 1)]
    assert abs(hubbard_bonds_ground_energy(n, chain, 3, 3, 1.0, 4.0) - hubbard_ground_energy(n, 3, 3, 1.0, 4.0)) < 1e-12
    assert abs(hubbard_bonds_ground_energy(n, chain, 3, 3, 1.0, 0.0) - free_fermion_energy(n, 3, 3)) < 1e-12


def test_add_mpo_is_the_weighted_sum():
    n = 5
    a = tk_hip.chain_mpo("heisenberg", n)
    b = tk_hip.compile_mpo(tk_hip.heisenberg_opsum(tk_hip.cylinder_bonds(1, 5) + [(0, 2)], n, J=0.7), n)
    s = tk_hip.add_mpo(a, b, 0.5, -2.0)
    want = 0.5 * _dense_from_mpo(a) - 2.0 * _dense_from_mpo(b)
    assert np.abs(_dense_from_mpo(s) - want).max() < 1e-13
    for W, Wa, Wb in zip(s[1:], a[1:], b[1:]):
        assert W.w.shape[2] == Wa.w.shape[2] + Wb.w.shape[2]
        assert len({tk_hip._qkey(q) for q, _ in W.wl}) == len(W.wl)  # one sector per charge
    with pytest.raises(ValueError):
        tk_hip.add_mpo(a, a[:-1])
