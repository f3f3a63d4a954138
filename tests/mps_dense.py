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


def from_dense(sym, legs, dirs, flux, dense):
    """Inverse of oracle.to_dense: block data in block-table order."""
    offs = []
    for leg in legs:
        o, acc = {}, 0
        for q, d in oracle._sorted_leg(leg):
            o[q] = (acc, d)
            acc += d
        offs.append(o)
    keys, boffs, tot = oracle.block_table(sym, legs, dirs, flux)
    out = np.zeros(tot)
    for k, off in zip(keys, boffs):
        qs = oracle.unpack_key(sym, k, len(legs))
        sl = tuple(slice(offs[i][qs[i]][0], offs[i][qs[i]][0] + offs[i][qs[i]][1]) for i in range(len(legs)))
        blk = dense[sl]
        out[off:off + blk.size] = blk.ravel()
    return out


def _bond_union(a, b):
    """Direct sum of two bond bases: per charge the a-part, then the b-part."""
    da = {_key(q): int(d) for q, d in a}
    db = {_key(q): int(d) for q, d in b}
    qs = sorted(set(da) | set(db))
    bond = [(q, da.get(q, 0) + db.get(q, 0)) for q in qs]
    pa, pb, acc = [], [], 0
    for q in qs:
        pa += list(range(acc, acc + da.get(q, 0)))
        acc += da.get(q, 0)
        pb += list(range(acc, acc + db.get(q, 0)))
        acc += db.get(q, 0)
    return bond, np.array(pa, dtype=int), np.array(pb, dtype=int)


def mps_sum(sym, phys, a, b, cb=1.0):
    """(bonds, sites) of |a> + cb |b> as a direct-sum MPS (bond dimensions add;
    the end bonds, of dimension one, are shared)."""
    (ba, sa), (bb, sb) = a, b
    n = len(phys)
    bonds, pos = [], []
    for i in range(n + 1):
        if i in (0, n):
            bonds.append(list(ba[i]))
            pos.append((np.array([0]), np.array([0])))
        else:
            bond, pa, pb = _bond_union(ba[i], bb[i])
            bonds.append(bond)
            pos.append((pa, pb))
    sites = []
    for i in range(n):
        Aa = oracle.to_dense(sym, [phys[i], ba[i], ba[i + 1]], [OUT, OUT, IN], (0, 0), sa[i])
        Ab = oracle.to_dense(sym, [phys[i], bb[i], bb[i + 1]], [OUT, OUT, IN], (0, 0), sb[i])
        dl = sum(d for _, d in bonds[i])
        dr = sum(d for _, d in bonds[i + 1])
        S = np.zeros((Aa.shape[0], dl, dr))
        S[np.ix_(np.arange(Aa.shape[0]), pos[i][0], pos[i + 1][0])] += Aa
        S[np.ix_(np.arange(Ab.shape[0]), pos[i][1], pos[i + 1][1])] += (cb if i == 0 else 1.0) * Ab
        sites.append(from_dense(sym, [phys[i], bonds[i], bonds[i + 1]], [OUT, OUT, IN], (0, 0), S))
    return bonds, sites
