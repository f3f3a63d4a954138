"""The J1 cylinder lattice and its MPO (CPU only).

* snake order: the reference's own known answer (tk-dsl/src/lattice/square.rs:90-97);
* bond count of the open square lattice (square.rs:84-88) plus one wrap bond per row;
* the MPO contracted to a dense matrix equals the bond sum built directly from
  Kronecker products (a misplaced or dropped FSA transition fails it);
* the oracle's DMRG ground energy on a 4 x 3 cylinder equals exact diagonalisation.
"""
import numpy as np
import pytest

import oracle
import tk_hip
from ed import heisenberg_bonds_ground_energy


def test_snake_path_reference():
    assert tk_hip.snake_path(3, 2) == [0, 1, 2, 5, 4, 3]


def test_cylinder_bonds():
    # 3 x 3 open square: 12 bonds (square.rs:84-88); the cylinder adds one per row
    assert len(tk_hip.cylinder_bonds(3, 3)) == 12 + 3
    b = tk_hip.cylinder_bonds(6, 12)
    assert len(b) == 6 * 12 * 2 - 6  # every site: one bond around, one along (minus the last row)
    assert len(set(b)) == len(b)
    # in snake order the wrap bond of row y joins positions 6y and 6y + 5
    assert (0, 5) in b and (6, 11) in b
    # width 2: no wrap bond (it would repeat the row's own bond)
    assert len(tk_hip.cylinder_bonds(2, 3)) == 3 + 4


def _dense_from_bonds(n, bonds, J=1.0):
    ops = {"z": tk_hip.S_Z, "p": tk_hip.S_PLUS, "m": tk_hip.S_MINUS}

    def site_op(i, o):
        out = np.ones((1, 1))
        for k in range(n):
            out = np.kron(out, ops[o] if k == i else np.eye(2))
        return out

    H = np.zeros((1 << n, 1 << n))
    for i, j in bonds:
        H += J * (site_op(i, "z") @ site_op(j, "z")
                  + 0.5 * (site_op(i, "p") @ site_op(j, "m") + site_op(i, "m") @ site_op(j, "p")))
    return H


def _dense_from_mpo(mpo):
    # W[s_in][s_out][a][b]; result H[(out...), (in...)] with site 0 most significant
    M = np.ones((1, 1, 1))  # (out, in, bond)
    for W in mpo:
        d = W.w.shape[0]
        # M'[o o', i i', b] = sum_a M[o, i, a] W[i', o', a, b]
        M = np.einsum("xya,ioab->xoyib", M, W.w)
        M = M.reshape(M.shape[0] * d, M.shape[2] * d, -1)
    return M[:, :, 0]


@pytest.mark.parametrize("width,length", [(3, 2), (4, 2), (2, 3)])
def test_cylinder_mpo_dense_equals_bond_sum(width, length):
    n = width * length
    bonds = tk_hip.cylinder_bonds(width, length)
    mpo = tk_hip.cylinder_mpo(width, length)
    assert len(mpo) == n
    assert np.abs(_dense_from_mpo(mpo) - _dense_from_bonds(n, bonds)).max() < 1e-14
    # the FSA bond dims on every cut: 2 + 3 * (sources with a partner beyond the cut)
    for c in range(n - 1):
        k = sum(1 for i in range(c + 1) if any(a == i and b > c for a, b in bonds))
        assert mpo[c].w.shape[3] == 2 + 3 * k


def test_cylinder_mpo_charges_conserved():
    # every nonzero W element conserves 2*Sz: q(out) - q(in) = q(w_l) - q(w_r)
    for W in tk_hip.cylinder_mpo(4, 3)[1:-1]:
        ql = [q[0] for (q, d) in W.wl for _ in range(d)]
        qr = [q[0] for (q, d) in W.wr for _ in range(d)]
        qs = [1, -1]
        for si, so, a, b in zip(*np.nonzero(W.w)):
            assert qs[so] - qs[si] == ql[a] - qr[b]


def test_oracle_dmrg_cylinder_matches_ed():
    width, length = 4, 3
    n = width * length
    mpo = tk_hip.cylinder_mpo(width, length)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), 16, 0)
    d = oracle.Dmrg(mpo, sym, schedule=(16, 64), davidson_tol=1e-12, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    st = [d.sweep() for _ in range(5)]
    e_ed = heisenberg_bonds_ground_energy(n, tk_hip.cylinder_bonds(width, length))
    assert abs(st[-1]["energy"] - e_ed) < 1e-9, (st[-1]["energy"], e_ed)
