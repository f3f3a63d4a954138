"""Small H_eff problems shared by the CPU oracle tests and the GPU parity tests."""
import numpy as np

import oracle
import tk_hip

UU, U1 = tk_hip.SYM_U1U1, tk_hip.SYM_U1
IN, OUT = tk_hip.DIR_IN, tk_hip.DIR_OUT


def sorted_perm(sectors):
    """Caller full index -> QIndex-sorted full index."""
    secs = [((q, 0) if np.isscalar(q) else tuple(q), int(d)) for q, d in sectors]
    order = sorted(range(len(secs)), key=lambda i: secs[i][0])
    off, acc = {}, 0
    for i in order:
        off[i] = acc
        acc += secs[i][1]
    perm = []
    for i, (_, d) in enumerate(secs):
        perm += list(range(off[i], off[i] + d))
    return np.array(perm)


def sorted_dense_w(W):
    """W as a dense array indexed by sorted full indices."""
    pp, pl, pr = sorted_perm(W.phys), sorted_perm(W.wl), sorted_perm(W.wr)
    out = np.zeros_like(W.w)
    out[np.ix_(pp, pp, pl, pr)] = W.w
    return out


def env_dense(sym, bond, wq, flat):
    """(bra, w, ket) dense environment in sorted full-index order."""
    keys, offs, tot = oracle.block_table(sym, [bond, wq, bond], [IN, IN, OUT])
    lb = sorted(((q, 0) if np.isscalar(q) else tuple(q), d) for q, d in bond)
    lw = sorted(((q, 0) if np.isscalar(q) else tuple(q), d) for q, d in wq)

    def offmap(l):
        o, a = {}, 0
        for q, d in l:
            o[q] = (a, d)
            a += d
        return o, a

    ob, nb = offmap(lb)
    ow, nw = offmap(lw)
    E = np.zeros((nb, nw, nb))
    for k, off in zip(keys, offs):
        qs = oracle.unpack_key(sym, k, 3)
        (a0, d0), (a1, d1), (a2, d2) = ob[qs[0]], ow[qs[1]], ob[qs[2]]
        E[a0:a0 + d0, a1:a1 + d1, a2:a2 + d2] = flat[off:off + d0 * d1 * d2].reshape(d0, d1, d2)
    return E


def random_problem(kind, seed, bond_l, bond_r):
    """Hubbard (U1xU1) or Heisenberg (U1) H_eff inputs with N(0,1) environments."""
    if kind == "hubbard":
        W, sym = tk_hip.hubbard_w(1.0, 4.0), UU
    else:
        W, sym = tk_hip.heisenberg_w(1.0), U1
    _, _, lt = oracle.block_table(sym, [bond_l, W.wl, bond_l], [IN, IN, OUT])
    _, _, rt = oracle.block_table(sym, [bond_r, W.wr, bond_r], [IN, IN, OUT])
    _, _, pt = oracle.block_table(sym, [W.phys, W.phys, bond_l, bond_r], [OUT, OUT, OUT, IN])
    rng = np.random.default_rng(seed)
    return {"bond_l": bond_l, "bond_r": bond_r, "W": W, "sym": sym, "L": rng.standard_normal(lt),
            "R": rng.standard_normal(rt), "x": rng.standard_normal(pt)}


TINY_UU_L = [((0, 0), 2), ((1, 0), 1), ((0, 1), 2), ((1, 1), 1), ((-1, 0), 1)]
TINY_UU_R = [((1, 1), 2), ((2, 1), 1), ((1, 2), 2), ((2, 2), 1), ((0, 1), 1), ((1, 0), 2)]
TINY_U1_L = [(0, 2), (2, 3), (-2, 1), (4, 1)]
TINY_U1_R = [(0, 3), (2, 2), (-2, 2), (4, 1), (-4, 1)]


def einsum_heff(p, x_dense):
    """Independent dense contraction with numpy (all indices sorted)."""
    W = sorted_dense_w(p["W"])
    L = env_dense(p["sym"], p["bond_l"], p["W"].wl, p["L"])
    R = env_dense(p["sym"], p["bond_r"], p["W"].wr, p["R"])
    # y[so_i, so_j, a', b'] = L[a', wl, a] Wi[si, so_i, wl, wm] Wj[sj, so_j, wm, wr] R[b', wr, b] x[si, sj, a, b]
    t = np.einsum("pwa,ijab->wijpb", L, x_dense)
    t = np.einsum("wijpb,ikwm->mkjpb", t, W)
    t = np.einsum("mkjpb,jlmr->klrpb", t, W)
    return np.einsum("klrpb,qrb->klpq", t, R)
