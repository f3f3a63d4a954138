"""Pins the oracle's block-sparse H_eff and block_gemm (CPU only).

* block-sparse H_eff == the reference's dense loop (environments/mod.rs:505-599)
  restated verbatim, at tiny bond dimension;
* == an independent numpy einsum of L W W psi R;
* diag == the diagonal of the dense matrix;
* block_gemm == dense GEMM on the LCG goldens of tk-linalg/tests/block_gemm_realistic.rs.
"""
import numpy as np
import pytest

import oracle
from conftest import lcg_next
from heff_cases import (TINY_U1_L, TINY_U1_R, TINY_UU_L, TINY_UU_R, U1, UU, einsum_heff, random_problem)


def _dense(p, flat):
    W = p["W"]
    return oracle.psi_to_dense(p["sym"], W.phys, W.phys, p["bond_l"], p["bond_r"], flat)


@pytest.mark.parametrize("kind,bl,br", [("hubbard", TINY_UU_L, TINY_UU_R), ("heisenberg", TINY_U1_L, TINY_U1_R)])
@pytest.mark.parametrize("seed", [0, 1])
def test_block_sparse_matches_reference_dense_loop(kind, bl, br, seed):
    p = random_problem(kind, seed, bl, br)
    y = oracle.heff_apply(p, p["x"], nthreads=2)
    xd = _dense(p, p["x"])
    yd_ref = oracle.heff_dense_reference(p, xd.ravel()).reshape(xd.shape)
    yd = _dense(p, y)
    assert np.abs(yd - yd_ref).max() <= 1e-12 * max(1.0, np.abs(yd_ref).max())
    # the dense loop writes nothing outside the symmetry-allowed blocks
    assert np.allclose(oracle.dense_to_psi(p["sym"], p["W"].phys, p["W"].phys, bl, br, yd_ref), y, atol=1e-12)


@pytest.mark.parametrize("kind,bl,br", [("hubbard", TINY_UU_L, TINY_UU_R), ("heisenberg", TINY_U1_L, TINY_U1_R)])
def test_block_sparse_matches_einsum(kind, bl, br):
    p = random_problem(kind, 5, bl, br)
    y = oracle.heff_apply(p, p["x"])
    ye = einsum_heff(p, _dense(p, p["x"]))
    assert np.abs(_dense(p, y) - ye).max() <= 1e-12 * np.abs(ye).max()


@pytest.mark.parametrize("kind,bl,br", [("hubbard", TINY_UU_L, TINY_UU_R), ("heisenberg", TINY_U1_L, TINY_U1_R)])
def test_diag_matches_dense(kind, bl, br):
    p = random_problem(kind, 3, bl, br)
    n = len(p["x"])
    d = oracle.heff_diag(p)
    dd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        dd[i] = oracle.heff_apply(p, e, nthreads=1)[i]
    assert np.abs(d - dd).max() <= 1e-12 * np.abs(dd).max()


def test_heisenberg_two_site_singlet():
    """Two-site chain with boundary MPOs: H_eff is H itself; lowest eigenvalue -3/4."""
    import tk_hip
    Wl, Wr = tk_hip.heisenberg_w(1.0, "left"), tk_hip.heisenberg_w(1.0, "right")
    triv = [(0, 1)]
    p = {"bond_l": triv, "bond_r": triv, "sym": U1, "L": np.ones(1), "R": np.ones(1)}
    _, _, pt = oracle.block_table(U1, [Wl.phys, Wr.phys, triv, triv], [1, 1, 1, 0])
    p["x"] = np.zeros(pt)
    H = np.zeros((pt, pt))
    for i in range(pt):
        e = np.zeros(pt)
        e[i] = 1.0
        H[:, i] = oracle.heff_apply(p, e, Wi=Wl, Wj=Wr)
    assert np.allclose(H, H.T)
    assert np.isclose(np.linalg.eigvalsh(H).min(), -0.75, atol=1e-14)


def _lcg_matrix(rows, cols, flux, seed):
    """build_test_matrix of block_gemm_realistic.rs:18-47: LCG fill per allowed
    block in (row QIndex, col QIndex) order; returned in key order with the dense form."""
    rs, cs = sorted(rows), sorted(cols)
    state = seed
    blocks = {}
    for rq, rd in rs:
        for cq, cd in cs:
            if rq - cq != flux:
                continue
            vals = []
            for _ in range(rd * cd):
                state, v = lcg_next(state)
                vals.append(v)
            blocks[oracle.pack_key(U1, [rq, cq])] = np.array(vals).reshape(rd, cd)
    flat = np.concatenate([blocks[k].ravel() for k in sorted(blocks)]) if blocks else np.zeros(0)
    ro = {q: o for (q, _), o in zip(rs, np.cumsum([0] + [d for _, d in rs])[:-1])}
    co = {q: o for (q, _), o in zip(cs, np.cumsum([0] + [d for _, d in cs])[:-1])}
    dense = np.zeros((sum(d for _, d in rs), sum(d for _, d in cs)))
    for k, b in blocks.items():
        (rq, _), (cq, _) = oracle.unpack_key(U1, k, 2)
        dense[ro[rq]:ro[rq] + b.shape[0], co[cq]:co[cq] + b.shape[1]] = b
    return flat, dense


def _to_dense(keys, offs, data, rows, cols):
    rs, cs = sorted(rows), sorted(cols)
    ro = {q: o for (q, _), o in zip(rs, np.cumsum([0] + [d for _, d in rs])[:-1])}
    co = {q: o for (q, _), o in zip(cs, np.cumsum([0] + [d for _, d in cs])[:-1])}
    rd, cd = dict(rs), dict(cs)
    dense = np.zeros((sum(rd.values()), sum(cd.values())))
    for k, o in zip(keys, offs):
        (rq, _), (cq, _) = oracle.unpack_key(U1, k, 2)
        m, n = rd[rq], cd[cq]
        dense[ro[rq]:ro[rq] + m, co[cq]:co[cq] + n] = data[o:o + m * n].reshape(m, n)
    return dense


@pytest.mark.parametrize("rows,cols,fa,fb,sa,sb", [
    ([(-1, 4), (0, 8), (1, 4)], [(-1, 3), (0, 6), (1, 3)], 0, 0, 42, 137),  # block_gemm_matches_dense_reference
    ([(0, 3), (1, 5)], [(-1, 2), (0, 4)], 1, -1, 99, 200),                   # block_gemm_nonzero_flux
])
def test_block_gemm_lcg_golden(rows, cols, fa, fb, sa, sb):
    a_flat, a_dense = _lcg_matrix(rows, cols, fa, sa)
    b_flat, b_dense = _lcg_matrix(cols, rows, fb, sb)
    keys, offs, c = oracle.block_gemm(U1, rows, cols, (fa, 0), a_flat, cols, rows, (fb, 0), b_flat)
    # output flux and every output sector obeys q_row - q_col = flux (faer.rs:398, tasks.rs:57-75)
    for k in keys:
        (rq, _), (cq, _) = oracle.unpack_key(U1, k, 2)
        assert rq - cq == fa + fb
    assert list(keys) == sorted(int(k) for k in keys)
    got = _to_dense(keys, offs, c, rows, rows)
    assert np.abs(got - a_dense @ b_dense).max() < 1e-10


def test_block_gemm_sector_count_bounded():
    # block_gemm_realistic.rs:166-196: 5 sectors -> at most 5 output sectors
    q = [(i, 2) for i in range(5)]
    keys, _, _ = oracle.block_gemm(U1, q, q, (0, 0), np.zeros(20), q, q, (0, 0), np.zeros(20))
    assert len(keys) <= 5
