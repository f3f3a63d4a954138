"""Exact diagonalisation of small open chains, built from bit strings with no
reference to any MPO (test infrastructure for pinning DMRG energies)."""
import itertools
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def heisenberg_ground_energy(n, J=1.0, twice_sz=0):
    """H = J sum_i S_i . S_{i+1}, open chain, sector 2*Sz = twice_sz."""
    n_up = (n + twice_sz) // 2
    states = [s for s in range(1 << n) if bin(s).count("1") == n_up]
    index = {s: k for k, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for k, s in enumerate(states):
        diag = 0.0
        for i in range(n - 1):
            a, b = (s >> i) & 1, (s >> (i + 1)) & 1
            diag += J * (0.25 if a == b else -0.25)
            if a != b:
                t = s ^ (1 << i) ^ (1 << (i + 1))
                rows.append(index[t])
                cols.append(k)
                vals.append(0.5 * J)
        rows.append(k)
        cols.append(k)
        vals.append(diag)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states), len(states)))
    if len(states) <= 400:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spl.eigsh(H, k=1, which="SA", tol=1e-14)[0][0])


def _hops(n, nparts):
    confs = [sum(1 << i for i in c) for c in itertools.combinations(range(n), nparts)]
    index = {c: k for k, c in enumerate(confs)}
    hops = []  # (k_from, k_to) for c+_i c_{i+1} and its conjugate; adjacent orbitals carry no sign
    for k, c in enumerate(confs):
        for i in range(n - 1):
            a, b = (c >> i) & 1, (c >> (i + 1)) & 1
            if a != b:
                hops.append((k, index[c ^ (1 << i) ^ (1 << (i + 1))]))
    return confs, hops


def hubbard_ground_energy(n, n_up, n_dn, t=1.0, U=4.0):
    """H = -t sum (c+_{i s} c_{i+1 s} + h.c.) + U sum n_up n_dn, open chain."""
    cu, hu = _hops(n, n_up)
    cd, hd = _hops(n, n_dn)
    nu, nd = len(cu), len(cd)
    Tu = sp.csr_matrix(([-t] * len(hu), ([b for _, b in hu], [a for a, _ in hu])), shape=(nu, nu))
    Td = sp.csr_matrix(([-t] * len(hd), ([b for _, b in hd], [a for a, _ in hd])), shape=(nd, nd))
    # spin-resolved ordering (all up orbitals, then all down): same-spin hops
    # between neighbouring sites cross no other orbital
    H = sp.kron(Tu, sp.identity(nd)) + sp.kron(sp.identity(nu), Td)
    dbl = np.array([bin(a & b).count("1") for a in cu for b in cd], dtype=np.float64)
    H = H + U * sp.diags(dbl)
    if nu * nd <= 500:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spl.eigsh(H.tocsr(), k=1, which="SA", tol=1e-14)[0][0])


def free_fermion_energy(n, n_up, n_dn, t=1.0):
    """U = 0 open chain: single-particle levels -2t cos(k pi / (n+1)), k = 1..n."""
    lev = sorted(-2 * t * math.cos(k * math.pi / (n + 1)) for k in range(1, n + 1))
    return sum(lev[:n_up]) + sum(lev[:n_dn])


def heisenberg_bonds_ground_energy(n, bonds, J=1.0, twice_sz=0):
    """H = J sum_{(i,j) in bonds} S_i . S_j on n spins, sector 2*Sz = twice_sz."""
    n_up = (n + twice_sz) // 2
    states = [s for s in range(1 << n) if bin(s).count("1") == n_up]
    index = {s: k for k, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for k, s in enumerate(states):
        diag = 0.0
        for i, j in bonds:
            a, b = (s >> i) & 1, (s >> j) & 1
            diag += J * (0.25 if a == b else -0.25)
            if a != b:
                rows.append(index[s ^ (1 << i) ^ (1 << j)])
                cols.append(k)
                vals.append(0.5 * J)
        rows.append(k)
        cols.append(k)
        vals.append(diag)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states), len(states)))
    if len(states) <= 400:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spl.eigsh(H, k=1, which="SA", tol=1e-14)[0][0])


def hubbard_bonds_ground_energy(n, bonds, n_up, n_dn, t=1.0, U=4.0):
    """H = -t sum_{(i,j) in bonds, s} (c+_{is} c_{js} + h.c.) + U sum n_up n_dn on
    n sites, spin-resolved orbital order (all up orbitals, then all down): a
    same-spin hop between i and j crosses the occupied orbitals of that spin
    strictly between them."""
    def sector(nparts):
        confs = [sum(1 << i for i in c) for c in itertools.combinations(range(n), nparts)]
        index = {c: k for k, c in enumerate(confs)}
        rows, cols, vals = [], [], []
        for k, c in enumerate(confs):
            for i, j in bonds:
                lo, hi = min(i, j), max(i, j)
                between = bin(c & (((1 << hi) - 1) ^ ((1 << (lo + 1)) - 1))).count("1")
                sgn = -1.0 if between % 2 else 1.0
                for a, b in ((i, j), (j, i)):  # c+_a c_b
                    if (c >> b) & 1 and not (c >> a) & 1:
                        rows.append(index[c ^ (1 << a) ^ (1 << b)])
                        cols.append(k)
                        vals.append(-t * sgn)
        return confs, sp.csr_matrix((vals, (rows, cols)), shape=(len(confs), len(confs)))

    cu, Tu = sector(n_up)
    cd, Td = sector(n_dn)
    H = sp.kron(Tu, sp.identity(len(cd))) + sp.kron(sp.identity(len(cu)), Td)
    dbl = np.array([bin(a & b).count("1") for a in cu for b in cd], dtype=np.float64)
    H = H + U * sp.diags(dbl)
    if H.shape[0] <= 500:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spl.eigsh(H.tocsr(), k=1, which="SA", tol=1e-14)[0][0])
