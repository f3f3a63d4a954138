"""Two-site DMRG with every contraction on the GPU, through the C ABI.

Parity bars: energies within 1e-9 of exact diagonalisation on chains ED can
hold; per-sweep energies and bond dimensions of the FP64 GPU path equal the FP64
oracle's on the same seeded MPS (|dE| <= 1e-10); the mixed-precision paths end
within 1e-10 of the FP64 path.
"""
import numpy as np
import pytest

import oracle
import tk_hip
from ed import (free_fermion_energy, heisenberg_bonds_ground_energy, heisenberg_ground_energy,
                hubbard_bonds_ground_energy, hubbard_ground_energy)

pytestmark = pytest.mark.gpu


def _setup(kind, n, total, d0=16, seed=0, **mp):
    mpo = tk_hip.chain_mpo(kind, n, **mp)
    sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, d0, seed)
    return mpo, sym, bonds, data


def _gpu(ctx, mpo, sym, bonds, data, sched, sweeps, prec="fp64", tol=1e-12, cutoff=1e-12, eigensolver="davidson"):
    d = tk_hip.Dmrg(ctx, mpo, sym, prec=prec, schedule=sched, davidson_tol=tol, svd_cutoff=cutoff,
                    eigensolver=eigensolver)
    d.set_state(bonds, data)
    d.prepare()
    st = [d.sweep() for _ in range(sweeps)]
    return d, st


def test_gpu_dmrg_heisenberg_two_sites(ctx):
    print("case dmrg heisenberg n=2", flush=True)
    mpo, sym, bonds, data = _setup("heisenberg", 2, (0, 0))
    d, st = _gpu(ctx, mpo, sym, bonds, data, (4,), 1)
    assert abs(st[-1]["energy"] + 0.75) < 1e-12


def test_gpu_dmrg_heisenberg_matches_ed(ctx):
    print("case dmrg heisenberg n=10", flush=True)
    n = 10
    mpo, sym, bonds, data = _setup("heisenberg", n, (0, 0))
    d, st = _gpu(ctx, mpo, sym, bonds, data, (32,), 4)
    assert abs(st[-1]["energy"] - heisenberg_ground_energy(n)) < 1e-9


def test_gpu_dmrg_hubbard_matches_ed(ctx):
    print("case dmrg hubbard n=6", flush=True)
    n = 6
    mpo, sym, bonds, data = _setup("hubbard", n, (3, 3), t=1.0, U=4.0)
    d, st = _gpu(ctx, mpo, sym, bonds, data, (64,), 4)
    assert abs(st[-1]["energy"] - hubbard_ground_energy(n, 3, 3, 1.0, 4.0)) < 1e-9
    mpo, sym, bonds, data = _setup("hubbard", n, (3, 3), t=1.0, U=0.0)
    d, st = _gpu(ctx, mpo, sym, bonds, data, (64,), 4)
    assert abs(st[-1]["energy"] - free_fermion_energy(n, 3, 3)) < 1e-9


def test_gpu_dmrg_lanczos_matches_ed_and_davidson(ctx):
    """The sweep with the Lanczos solver (eigensolver/mod.rs:57-297) reaches the
    same ground state as with Davidson and as exact diagonalisation."""
    print("case dmrg lanczos heisenberg n=10, hubbard n=6", flush=True)
    n = 10
    mpo, sym, bonds, data = _setup("heisenberg", n, (0, 0))
    d, st = _gpu(ctx, mpo, sym, bonds, data, (32,), 4, eigensolver="lanczos")
    e_l = st[-1]["energy"]
    assert abs(e_l - heisenberg_ground_energy(n)) < 1e-9
    d, st = _gpu(ctx, mpo, sym, bonds, data, (32,), 4)
    assert abs(e_l - st[-1]["energy"]) < 1e-10
    mpo, sym, bonds, data = _setup("hubbard", 6, (3, 3), t=1.0, U=4.0)
    d, st = _gpu(ctx, mpo, sym, bonds, data, (64,), 4, eigensolver="lanczos")
    assert abs(st[-1]["energy"] - hubbard_ground_energy(6, 3, 3, 1.0, 4.0)) < 1e-9


def _oracle(mpo, sym, bonds, data, sched, sweeps, tol=1e-12, cutoff=1e-12):
    o = oracle.Dmrg(mpo, sym, schedule=sched, davidson_tol=tol, svd_cutoff=cutoff)
    o.set_state(bonds, data)
    o.prepare()
    return o, [o.sweep() for _ in range(sweeps)]


@pytest.mark.parametrize("kind,n,total,sched,sweeps", [
    ("heisenberg", 32, (0, 0), (64,), 4),      # the CPU-runnable workload: L=32, D=64, 4 sweeps
    ("hubbard", 12, (6, 6), (24, 48), 3),
])
def test_gpu_dmrg_matches_oracle(ctx, kind, n, total, sched, sweeps):
    print(f"case dmrg vs oracle {kind} n={n}", flush=True)
    mp = {"t": 1.0, "U": 4.0} if kind == "hubbard" else {}
    mpo, sym, bonds, data = _setup(kind, n, total, **mp)
    d, st = _gpu(ctx, mpo, sym, bonds, data, sched, sweeps, cutoff=1e-10)
    o, so = _oracle(mpo, sym, bonds, data, sched, sweeps, cutoff=1e-10)
    for a, b in zip(st, so):
        assert abs(a["energy"] - b["energy"]) <= 1e-10, (a["energy"], b["energy"])
        assert a["max_bond_dim"] == int(b["max_bond_dim"])
        assert abs(a["max_truncation_error"] - b["max_truncation_error"]) <= 1e-9
    for i in range(n + 1):
        assert d.bond(i) == o.bond(i), i


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_gpu_dmrg_mixed_precision_matches_fp64(ctx, prec):
    print(f"case dmrg mixed {prec}", flush=True)
    mpo, sym, bonds, data = _setup("hubbard", 12, (6, 6), t=1.0, U=4.0)
    d64, s64 = _gpu(ctx, mpo, sym, bonds, data, (24, 48), 3, tol=1e-10)
    dlo, slo = _gpu(ctx, mpo, sym, bonds, data, (24, 48), 3, prec=prec, tol=1e-10)
    assert abs(slo[-1]["energy"] - s64[-1]["energy"]) <= 1e-10
    # low precision expands the basis of the cold (first-sweep) solves; warm
    # solves multiply their late corrections (residual below 1000 tol) in low
    # precision
    assert slo[0]["matvecs_fp64"] < slo[0]["matvecs"]
    assert slo[-1]["matvecs_fp64"] < slo[-1]["matvecs"]
    assert s64[-1]["matvecs_fp64"] == s64[-1]["matvecs"]


def test_gpu_dmrg_cylinder_matches_ed(ctx):
    """J1 Heisenberg on a 4 x 3 cylinder in snake order (MPO bond dims up to 14)."""
    print("case dmrg cylinder 4x3", flush=True)
    width, length = 4, 3
    mpo = tk_hip.cylinder_mpo(width, length)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), 16, 0)
    d, st = _gpu(ctx, mpo, sym, bonds, data, (16, 64), 5)
    e_ed = heisenberg_bonds_ground_energy(width * length, tk_hip.cylinder_bonds(width, length))
    assert abs(st[-1]["energy"] - e_ed) < 1e-9


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_gpu_dmrg_cylinder_mixed_matches_fp64(ctx, prec):
    """6 x 3 cylinder (MPO bond dim 20) truncated at D = 64: the mixed path
    ends within 1e-10 of the FP64 path with the same truncation."""
    print(f"case dmrg cylinder 6x3 mixed {prec}", flush=True)
    mpo = tk_hip.cylinder_mpo(6, 3)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), 16, 0)
    d64, s64 = _gpu(ctx, mpo, sym, bonds, data, (32, 64), 4, tol=1e-10)
    dlo, slo = _gpu(ctx, mpo, sym, bonds, data, (32, 64), 4, prec=prec, tol=1e-10)
    assert abs(slo[-1]["energy"] - s64[-1]["energy"]) <= 1e-10
    assert abs(slo[-1]["max_truncation_error"] - s64[-1]["max_truncation_error"]) <= 1e-9


def test_gpu_dmrg_checkpoint_resume(ctx, tmp_path):
    """Two sweeps, checkpoint to disk, a new engine resumes: the saved MPS
    comes back bit-exact and two more sweeps reach the ED energy."""
    print("case dmrg checkpoint", flush=True)
    n = 10
    mpo, sym, bonds, data = _setup("heisenberg", n, (0, 0))
    d1, s1 = _gpu(ctx, mpo, sym, bonds, data, (16, 32), 2)
    cp = d1.checkpoint(2, s1[-1]["energy"], {"max_bond_dim": s1[-1]["max_bond_dim"]})
    path = tmp_path / "dmrg.npz"
    cp.write_to_file(path)
    back = tk_hip.DMRGCheckpoint.read_from_file(path)
    d2 = tk_hip.Dmrg(ctx, mpo, sym, schedule=(32,), davidson_tol=1e-12, svd_cutoff=1e-12)
    d2.restore(back, prepare=False)
    for i in range(n):
        assert np.array_equal(d2.site(i), d1.site(i))
    for i in range(n + 1):
        assert d2.bond(i) == d1.bond(i)
    d2.prepare()
    st = [d2.sweep() for _ in range(2)]
    assert abs(st[-1]["energy"] - heisenberg_ground_energy(n)) < 1e-9


def test_gpu_dmrg_compiled_hubbard_ladder_matches_ed(ctx):
    """The OpSum compiler's Hubbard MPO on a 3 x 2 ladder in snake order
    (hops across other sites carry the Jordan-Wigner string) on the GPU."""
    print("case dmrg compiled hubbard ladder", flush=True)
    n = 6
    bonds = tk_hip.cylinder_bonds(2, 3)
    mpo = tk_hip.compile_mpo(tk_hip.hubbard_opsum(bonds, n, 1.0, 4.0), n, "electron")
    sym = tk_hip.SYM_U1U1
    b, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (3, 3), 16, 0)
    d, st = _gpu(ctx, mpo, sym, b, data, (64, 128), 5)
    assert abs(st[-1]["energy"] - hubbard_bonds_ground_energy(n, bonds, 3, 3, 1.0, 4.0)) < 1e-9
