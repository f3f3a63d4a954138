"""SVD compression of MPOs (compress_mpo) on the device.

Known answers from the operator Schmidt rank of the Hamiltonians:
* H + H of the Heisenberg chain (bond 10 as a direct sum) compresses to the
  chain's own rank, bonds 1, 4, 5, ..., 5, 4, 1, and densifies to 2H;
* the Hubbard chain's H + H compresses to bond 6 on every inner cut;
* exponentially decaying Sz Sz couplings sum_{i<j} exp(-(j-i)/xi) Sz_i Sz_j,
  which the finite-state compiler opens one channel per source for (bond
  2 + open sources), have rank 3 on every bulk cut (2 next to the ends);
* an already minimal MPO (the compiled 3 x 3 cylinder) keeps its operator.
Truncating below the rank gives an error bounded by the returned
error_bound, and validation rejects it at compression_tol.  A DMRG run on
the compressed 2H gives twice the exact ground energy.
"""
import numpy as np
import pytest

import tk_hip
from ed import heisenberg_ground_energy
from test_cylinder_mpo import _dense_from_mpo

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _exp_szsz(n, xi=2.0):
    os_ = tk_hip.OpSum(n)
    for i in range(n):
        for j in range(i + 1, n):
            os_.add(np.exp(-(j - i) / xi), ("Sz", i), ("Sz", j))
    return tk_hip.compile_mpo(os_, n)


def _charges_conserved(mpo):
    for W in mpo:
        _, blocks = tk_hip._bond_blocks(W, "right")  # raises when broken
        assert blocks


def test_heisenberg_sum_compresses_to_rank(ctx):
    print("case compress heisenberg 2H", flush=True)
    n = 8
    h = tk_hip.chain_mpo("heisenberg", n)
    s = tk_hip.add_mpo(h, h)
    assert [W.w.shape[3] for W in s[:-1]] == [10] * (n - 1)
    c, info = tk_hip.compress_mpo(ctx, s)
    assert info["bond_dims"] == [1, 4] + [5] * (n - 3) + [4, 1]
    assert info["error_bound"] < 1e-10
    assert _rel(_dense_from_mpo(c), 2 * _dense_from_mpo(h)) < 1e-12
    _charges_conserved(c)


def test_hubbard_sum_compresses_to_rank(ctx):
    print("case compress hubbard 2H", flush=True)
    n = 5
    h = tk_hip.chain_mpo("hubbard", n, t=1.0, U=4.0)
    c, info = tk_hip.compress_mpo(ctx, tk_hip.add_mpo(h, h, 0.5, 1.5))
    assert info["bond_dims"] == [1] + [6] * (n - 1) + [1]
    assert _rel(_dense_from_mpo(c), 2 * _dense_from_mpo(h)) < 1e-12
    _charges_conserved(c)


def test_exponential_couplings_compress_to_rank_three(ctx):
    print("case compress exp SzSz", flush=True)
    n = 10
    m = _exp_szsz(n)
    assert max(W.w.shape[3] for W in m) > 3
    c, info = tk_hip.compress_mpo(ctx, m)
    assert info["bond_dims"] == [1, 2] + [3] * (n - 3) + [2, 1]
    assert _rel(_dense_from_mpo(c), _dense_from_mpo(m)) < 1e-12


def test_minimal_cylinder_keeps_operator(ctx):
    print("case compress cylinder 3x3", flush=True)
    b = tk_hip.cylinder_bonds(3, 3)
    m = tk_hip.compile_mpo(tk_hip.heisenberg_opsum(b, 9), 9)
    c, info = tk_hip.compress_mpo(ctx, m)
    orig = [1] + [W.w.shape[3] for W in m[:-1]] + [1]
    assert all(a <= o for a, o in zip(info["bond_dims"], orig))
    assert _rel(_dense_from_mpo(c), _dense_from_mpo(m)) < 1e-12


def test_truncated_compression_error_bound(ctx):
    print("case compress truncated", flush=True)
    n = 8
    m = _exp_szsz(n, xi=1.0)
    # couplings through a second channel: sum of two decay lengths has rank 4
    m2 = tk_hip.add_mpo(m, _exp_szsz(n, xi=3.0))
    full, info_full = tk_hip.compress_mpo(ctx, m2)
    assert max(info_full["bond_dims"]) == 4
    with pytest.raises(ValueError):
        tk_hip.compress_mpo(ctx, m2, max_bond_dim=3)
    c, info = tk_hip.compress_mpo(ctx, m2, max_bond_dim=3, validate_compression=False)
    assert max(info["bond_dims"]) == 3
    h, hc = _dense_from_mpo(m2), _dense_from_mpo(c)
    err = np.linalg.norm(hc - h) / np.linalg.norm(h)
    assert 0 < err <= info["error_bound"] * (1 + 1e-9)
    assert info["error_bound"] > 1e-8


def test_dmrg_on_compressed_sum(ctx):
    print("case compress dmrg", flush=True)
    n = 10
    h = tk_hip.chain_mpo("heisenberg", n)
    c, _ = tk_hip.compress_mpo(ctx, tk_hip.add_mpo(h, h))
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [W.phys for W in c], (0, 0), 8, 0)
    d = tk_hip.Dmrg(ctx, c, sym, schedule=(16, 32, 64), davidson_tol=1e-12, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    e = [d.sweep()["energy"] for _ in range(6)][-1]
    d.close()
    assert abs(e - 2 * heisenberg_ground_energy(n)) < 1e-9
