"""Pins the oracle's eigensolver, SVD and DMRG (CPU only).

* Davidson / Jacobi known answers of tk-dmrg/src/eigensolver/mod.rs:656-801;
* truncation known answers of tk-dmrg/src/truncation/mod.rs:174-259;
* one-sided Jacobi SVD vs LAPACK (numpy) and the faer-test matrices;
* full DMRG energies vs exact diagonalisation and the U=0 closed form.
"""
import math

import numpy as np
import pytest

import oracle
import tk_hip
from ed import free_fermion_energy, heisenberg_ground_energy, hubbard_ground_energy


def test_jacobi_small_three_by_three():
    # symmetric_eigen_small_test: [[2,-1,0],[-1,2,-1],[0,-1,2]] -> 2 - sqrt(2)
    ev, _ = oracle.jacobi_small(np.array([[2.0, -1, 0], [-1, 2, -1], [0, -1, 2]]))
    assert abs(min(ev) - (2 - math.sqrt(2))) < 1e-10
    assert np.allclose(sorted(ev), [2 - math.sqrt(2), 2, 2 + math.sqrt(2)], atol=1e-12)


def test_davidson_known_answers():
    # davidson_diagonal_matrix: diag(3,1,4,1,5) with the preconditioner -> 1
    d = np.array([3.0, 1, 4, 1, 5])
    e, _, _ = oracle.davidson_dense(np.diag(d), diag=d)
    assert abs(e - 1.0) < 1e-6
    # davidson_tridiagonal_2x2: [[2,-1],[-1,2]], diag (2,2), tol 1e-8 -> 1
    e, _, _ = oracle.davidson_dense(np.array([[2.0, -1], [-1, 2]]), diag=np.array([2.0, 2.0]), tol=1e-8)
    assert abs(e - 1.0) < 1e-6
    # davidson_without_preconditioner: diag(5,2,8,1,3) -> 1
    e, _, _ = oracle.davidson_dense(np.diag([5.0, 2, 8, 1, 3]), tol=1e-8)
    assert abs(e - 1.0) < 1e-6


def test_davidson_random_matches_lapack():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((60, 60))
    h = a + a.T
    e, x, _ = oracle.davidson_dense(h, diag=np.diag(h).copy(), tol=1e-11)
    w, v = np.linalg.eigh(h)
    assert abs(e - w[0]) < 1e-10
    assert np.linalg.norm(h @ x - e * x) < 1e-9


@pytest.mark.parametrize("shape", [(3, 2), (2, 3), (7, 7), (12, 5), (5, 12), (40, 33)])
def test_svd_matches_lapack(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    u, s, v = oracle.svd(a)
    assert np.allclose(s, np.linalg.svd(a, compute_uv=False), atol=1e-13)
    assert np.abs(u @ np.diag(s) @ v.T - a).max() < 1e-12
    assert np.abs(u.T @ u - np.eye(len(s))).max() < 1e-12
    assert np.abs(v.T @ v - np.eye(len(s))).max() < 1e-12


def test_svd_faer_matrix():
    # faer.rs:795-844: [[1,2],[3,4],[5,6]] reconstructs; sigma descending
    a = np.array([[1.0, 2], [3, 4], [5, 6]])
    u, s, v = oracle.svd(a)
    assert s[0] >= s[1] > 0
    assert np.abs(u @ np.diag(s) @ v.T - a).max() < 1e-12


def _run_oracle(kind, n, sched, sweeps, total, d0=16, seed=0, tol=1e-12, **mp):
    mpo = tk_hip.chain_mpo(kind, n, **mp)
    sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, d0, seed)
    d = oracle.Dmrg(mpo, sym, schedule=sched, davidson_tol=tol, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    return d, [d.sweep() for _ in range(sweeps)]


def test_dmrg_heisenberg_two_sites():
    d, st = _run_oracle("heisenberg", 2, (4,), 1, (0, 0))
    assert abs(st[-1]["energy"] + 0.75) < 1e-12


def test_dmrg_heisenberg_matches_ed():
    n = 10
    d, st = _run_oracle("heisenberg", n, (32,), 4, (0, 0))
    e_ed = heisenberg_ground_energy(n)
    assert abs(st[-1]["energy"] - e_ed) < 1e-9, (st[-1]["energy"], e_ed)


def test_dmrg_hubbard_matches_ed():
    n = 6
    d, st = _run_oracle("hubbard", n, (64,), 4, (3, 3), t=1.0, U=4.0)
    e_ed = hubbard_ground_energy(n, 3, 3, 1.0, 4.0)
    assert abs(st[-1]["energy"] - e_ed) < 1e-9, (st[-1]["energy"], e_ed)


def test_dmrg_free_fermions_closed_form():
    n = 6
    d, st = _run_oracle("hubbard", n, (64,), 4, (3, 3), t=1.0, U=0.0)
    assert abs(st[-1]["energy"] - free_fermion_energy(n, 3, 3)) < 1e-9


def test_dmrg_hubbard_away_from_half_filling():
    n = 6
    d, st = _run_oracle("hubbard", n, (64,), 4, (2, 3), t=1.0, U=4.0)
    assert abs(st[-1]["energy"] - hubbard_ground_energy(n, 2, 3, 1.0, 4.0)) < 1e-9


def test_ed_closed_form_cross_check():
    # the ED itself against the U=0 closed form
    assert abs(hubbard_ground_energy(6, 3, 3, 1.0, 0.0) - free_fermion_energy(6, 3, 3)) < 1e-10
    # two-site Heisenberg singlet
    assert abs(heisenberg_ground_energy(2) + 0.75) < 1e-14
