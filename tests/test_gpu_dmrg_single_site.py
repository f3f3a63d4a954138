"""Single-site DMRG sweeps on the GPU (dmrg_step_single_site,
tk-dmrg/src/sweep/mod.rs:451-509, with the eigenvector written back).

The reference step leaves the MPS unchanged, so the sweep is pinned by
exact diagonalisation and the variational principle: with every bond at full
dimension the one-site space is complete and the sweeps reach the ED ground
energy (1e-9); at a truncated bond dimension they never raise the energy of
the two-site result and never change a bond dimension."""
import numpy as np
import pytest

import tk_hip
from ed import heisenberg_ground_energy, hubbard_ground_energy

pytestmark = pytest.mark.gpu


def _engine(ctx, kind, n, total, d0, sched, **mp):
    mpo = tk_hip.chain_mpo(kind, n, **mp)
    sym = tk_hip.SYM_U1 if kind == "heisenberg" else tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], total, d0, seed=5)
    d = tk_hip.Dmrg(ctx, mpo, sym, schedule=sched, davidson_tol=1e-12, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    return d


def test_gpu_single_site_full_bonds_reaches_ed(ctx):
    print("case single-site heisenberg n=8 full bonds", flush=True)
    n = 8
    d = _engine(ctx, "heisenberg", n, (0, 0), 256, (256,))
    dims0 = [sum(x for _, x in d.bond(i)) for i in range(n + 1)]
    e = [d.sweep_single_site()["energy"] for _ in range(6)]
    assert abs(e[-1] - heisenberg_ground_energy(n)) < 1e-9
    assert all(e[k + 1] <= e[k] + 1e-10 for k in range(len(e) - 1))
    assert all(a <= b for a, b in zip([sum(x for _, x in d.bond(i)) for i in range(n + 1)], dims0))


def test_gpu_single_site_after_two_site(ctx):
    print("case single-site hubbard n=8 D=24 after two-site", flush=True)
    n = 8
    d = _engine(ctx, "hubbard", n, (4, 4), 16, (24,), t=1.0, U=4.0)
    e2 = [d.sweep()["energy"] for _ in range(4)][-1]
    dims = [sum(x for _, x in d.bond(i)) for i in range(n + 1)]
    st = [d.sweep_single_site() for _ in range(2)]
    assert st[-1]["energy"] <= e2 + 1e-10
    assert st[-1]["energy"] >= hubbard_ground_energy(n, 4, 4, 1.0, 4.0) - 1e-10
    assert all(a <= b for a, b in zip([sum(x for _, x in d.bond(i)) for i in range(n + 1)], dims))
    assert st[-1]["max_truncation_error"] < 1e-20 and st[-1]["unconverged"] == 0


def test_gpu_single_site_converged_state_is_stationary(ctx):
    print("case single-site hubbard n=6 converged", flush=True)
    n = 6
    d = _engine(ctx, "hubbard", n, (3, 3), 64, (64,), t=1.0, U=4.0)
    for _ in range(4):
        d.sweep()
    e = d.sweep_single_site()["energy"]
    assert abs(e - hubbard_ground_energy(n, 3, 3, 1.0, 4.0)) < 1e-9
    steps = d.steps()
    assert len(steps) == 2 * (n - 1)
    assert np.ptp([s["energy"] for s in steps]) < 1e-9


def _dims(d, n):
    return [sum(x for _, x in d.bond(i)) for i in range(n + 1)]


@pytest.mark.parametrize("kind,n,total,sched,mp", [
    ("heisenberg", 10, (0, 0), (32,), {}),
    ("hubbard", 6, (3, 3), (64,), {"t": 1.0, "U": 4.0}),
])
def test_gpu_single_site_subspace_expansion_reaches_ed(ctx, kind, n, total, sched, mp):
    """From a small start (one state per allowed charge, at most two per
    sector) a plain single-site sweep can never grow a bond and stays above
    the ground state; with subspace expansion the bonds grow to what the
    schedule allows.  A fixed alpha leaves the energy about alpha^2 |E| above
    ED (the bases of the perturbed moves); the adaptive alpha falls tenfold
    per settled sweep and the sweeps reach the ED energy."""
    print(f"case single-site expansion {kind} n={n}", flush=True)
    e_ed = heisenberg_ground_energy(n) if kind == "heisenberg" else hubbard_ground_energy(n, *total, 1.0, 4.0)
    plain = _engine(ctx, kind, n, total, 2, sched, **mp)
    dims0 = _dims(plain, n)
    e_plain = [plain.sweep_single_site()["energy"] for _ in range(3)][-1]
    print(f"plain bonds {dims0} E - E_ED = {e_plain - e_ed:.3e}", flush=True)
    assert _dims(plain, n) == dims0
    assert e_plain > e_ed + 1e-4
    plain.close()
    fixed = _engine(ctx, kind, n, total, 2, sched, **mp)
    fixed.subspace_expansion(1e-4, adaptive=False)
    ef = [fixed.sweep_single_site()["energy"] for _ in range(4)]
    assert max(_dims(fixed, n)) > max(dims0)
    assert 0 <= ef[-1] - e_ed < 1e-6 and fixed.subspace_expansion_alpha() == 1e-4
    fixed.close()
    d = _engine(ctx, kind, n, total, 2, sched, **mp)
    d.subspace_expansion(1e-4)
    e = [d.sweep_single_site()["energy"] for _ in range(14)]
    print(f"expanded bonds {_dims(d, n)} alpha {d.subspace_expansion_alpha():.1e} "
          f"E - E_ED = {[f'{x - e_ed:.1e}' for x in e]}", flush=True)
    assert max(_dims(d, n)) > max(dims0)
    assert all(e[k + 1] <= e[k] + 1e-6 for k in range(len(e) - 1))
    assert d.subspace_expansion_alpha() < 1e-4
    assert abs(e[-1] - e_ed) < 1e-9
    d.close()


def test_gpu_single_site_expansion_cap_and_off(ctx):
    """max_new bounds the growth of a bond per centre move (two moves per
    sweep); alpha = 0 restores the bond-preserving sweep; a negative alpha is
    rejected."""
    print("case single-site expansion cap", flush=True)
    n = 10
    d = _engine(ctx, "heisenberg", n, (0, 0), 2, (32,))
    before = _dims(d, n)
    d.subspace_expansion(1e-4, max_new=2)
    d.sweep_single_site()
    after = _dims(d, n)
    assert all(a <= b + 4 for a, b in zip(after, before)) and max(after) > max(before)
    d.subspace_expansion(0.0)
    d.sweep_single_site()
    assert _dims(d, n) == after
    with pytest.raises(tk_hip.TkError) as err:
        d.subspace_expansion(-1.0)
    assert err.value.status == tk_hip.TK_ERR_INVALID_ARG
    d.close()
