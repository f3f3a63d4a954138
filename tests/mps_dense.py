"""Dense forms of small MPOs and MPS for tests (site 0 most significant).

Physical indices follow the sorted QIndex order of the block tables (the
order oracle.to_dense uses), so the dense MPO and the dense MPS share one basis."""
import numpy as np

import oracle

IN, OUT = 0, 1


def _key(q):
    return (int(q), 0) if np.isscalar(q) else tuple(int(c) for c in q)


def sorted_phys(phys):
    """(perm, charges): perm[caller full index] = sorted full index; charges[sorted index]."""
    secs = [(_key(q), int(d)) for q, d in phys]
    order = sorted(range(len(secs)), key=lambda i: secs[i][0])
    off, acc, charges = {}, 0, []
    for i in order:
        off[i] = acc
        acc += secs[i][1]
        charges += [secs[i][0]] * secs[i][1]
    perm = [off[i] + x for i, (_, d) in enumerate(secs) for x in range(d)]
    return np.array(perm), charges


def dense_mpo(mpo):
    """H[(out...), (in...)] over the sorted physical bases, and the charge of every basis state."""
    M = np.ones((1, 1, 1))
    charges = [(0, 0)]
    for W in mpo:
        p, ch = sorted_phys(W.phys)
        Ws = np.zeros_like(W.w)
        Ws[np.ix_(p, p)] = W.w
        d = Ws.shape[0]
        M = np.einsum("xya,ioab->xoyib", M, Ws).reshape(M.shape[0] * d, M.shape[1] * d, -1)
        charges = [(a[0] + b[0], a[1] + b[1]) for a in charges for b in ch]
    return M[:, :, 0], charges


def dense_mps(sym, mpo, bonds, sites):
    """Dense state of an MPS given its bonds and per-site block data."""
    psi = np.ones((1, 1))
    for i, W in enumerate(mpo):
        A = oracle.to_dense(sym, [W.phys, bonds[i], bonds[i + 1]], [OUT, OUT, IN], (0, 0), sites[i])
        psi = np.einsum("xa,sab->xsb", psi, A).reshape(psi.shape[0] * A.shape[0], A.shape[2])
    return psi[:, 0]
