"""FP64 CPU oracle for the DMFT pieces (numpy / scipy).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product
package).  Each function names the reference it restates and how it is
pinned:

* hyb_star (bath.rs:61-83): the reference's own known value at omega = 0,
  eta = 0.1 (bath.rs:180-195); hyb_chain: equals hyb_star of the same bath
  after the chain mapping (an exact identity of the Lanczos recursion) and the
  one-site closed form v^2 / (z - eps).
* lanczos_tridiagonalize (discretize.rs:54-274): the reference tests
  (discretize.rs:288-375: bath count, energies within the band, V_1 > 0) and
  exact recovery of a discrete bath (chain Delta == star Delta).
* uniform / linear_mix (bath.rs:37-55, 124-166): bath.rs:170-178 (-5 .. 5)
  and bath.rs:206-223 (-1.3, 1.3, 0.85).
* Broyden (mixing.rs:40-150): mixing.rs:176-229 (first step linear, second
  differs, reset after the history depth).
* aim_sector_hamiltonian / aim_ground_state / aim_green: exact
  diagonalisation with explicit Jordan-Wigner bit arithmetic (spin-orbitals
  ordered 0up, 0dn, 1up, 1dn, ...), independent of the MPO convention; pinned
  by the atomic limit G = (1/(z - U/2) + 1/(z + U/2)) / 2, the U = 0 closed
  form G = 1 / (z - eps_imp - Delta(z)) and the free-fermion ground-state
  energy (twice the sum of the negative one-particle levels).
* mlp_forward / mlp_grad: float64 restatement of the network; the gradient
  is pinned by central finite differences.
"""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------- baths
def hyb_star(eps, v, z):
    z = np.asarray(z, dtype=np.complex128)
    out = np.zeros(z.shape, dtype=np.complex128)
    for e, vk in zip(eps, v):
        out += (vk * vk) / (z - e)
    return out


def hyb_chain(eps, v, z):
    z = np.asarray(z, dtype=np.complex128)
    g = np.zeros(z.shape, dtype=np.complex128)
    for k in range(len(eps) - 1, -1, -1):
        g = (v[k] * v[k]) / (z - eps[k] - g)
    return g


def uniform(n_bath, bandwidth, v0):
    half = bandwidth / 2.0
    eps = [0.0] * n_bath if n_bath <= 1 else [-half + k / (n_bath - 1.0) * bandwidth for k in range(n_bath)]
    return np.array(eps, dtype=float), np.full(n_bath, float(v0))


def linear_mix(a, b, alpha):
    return (1 - alpha) * np.asarray(a[0]) + alpha * np.asarray(b[0]), (1 - alpha) * np.asarray(a[1]) + alpha * np.asarray(b[1])


def lanczos_points(x, w, n_bath):
    """The recursion of discretize.rs:114-233 on points x with weights w."""
    x = np.asarray(x, float)
    w = np.asarray(w, float)
    eps = np.zeros(n_bath)
    v = np.zeros(n_bath)
    if n_bath == 0 or x.size == 0:
        return eps, v
    W = w.sum()
    if W < np.finfo(float).eps:
        return eps, v
    v1 = math.sqrt(W)
    prev = np.zeros_like(x)
    cur = np.full_like(x, 1.0 / v1)
    steps = min(n_bath, x.size)
    al, be = [], [v1]
    for k in range(steps):
        a = float(np.sum(w * x * cur * cur))
        al.append(a)
        if k + 1 >= steps:
            break
        bk = be[k] if k > 0 else 0.0
        r = x * cur - a * cur - bk * prev
        b = math.sqrt(float(np.sum(w * r * r)))
        if b < 1e-15:
            be.append(0.0)
            break
        be.append(b)
        prev, cur = cur, r / b
    for k in range(n_bath):
        eps[k] = al[k] if k < len(al) else 0.0
        v[k] = be[k] if k < len(be) else 0.0
    return eps, v


class Broyden:
    """Restatement of BroydenState::update (mixing.rs:75-150)."""

    def __init__(self, depth):
        self.depth = depth
        self.prev_x = self.prev_r = self.J = None
        self.n_updates = 0

    def update(self, x, f, alpha):
        x = np.asarray(x, float)
        f = np.asarray(f, float)
        r = f - x
        n = x.size
        if self.prev_x is None or self.J.shape[0] != n or self.n_updates >= self.depth:
            self.n_updates = 0
            self.J = -alpha * np.eye(n)
            self.prev_x, self.prev_r = x.copy(), r.copy()
            return (1 - alpha) * x + alpha * f
        dx = x - self.prev_x
        df = r - self.prev_r
        jdf = self.J @ df
        num = dx - jdf
        dxj = dx @ self.J
        den = float(dxj @ df)
        if abs(den) > 1e-15:
            self.J = self.J + np.outer(num, dxj) / den
        out = x - self.J @ r
        self.prev_x, self.prev_r = x.copy(), r.copy()
        self.n_updates += 1
        return out


# ---------------------------------------------------------------- exact diagonalisation
def _popcount(x):
    return bin(x).count("1")


def sector_states(n_sites, n_up, n_dn):
    """Occupation bit patterns (bit 2i: site i up, bit 2i+1: site i down)."""
    out = []
    for s in range(1 << (2 * n_sites)):
        nu = sum((s >> (2 * i)) & 1 for i in range(n_sites))
        nd = sum((s >> (2 * i + 1)) & 1 for i in range(n_sites))
        if nu == n_up and nd == n_dn:
            out.append(s)
    return out


def _cdag_c(s, i, j):
    """c+_i c_j |s> for spin-orbitals i, j (Jordan-Wigner order 0, 1, ...):
    (sign, new state) or None."""
    if not (s >> j) & 1:
        return None
    sign = -1 if _popcount(s & ((1 << j) - 1)) % 2 else 1
    s2 = s & ~(1 << j)
    if (s2 >> i) & 1:
        return None
    sign *= -1 if _popcount(s2 & ((1 << i) - 1)) % 2 else 1
    return sign, s2 | (1 << i)


def aim_sector_hamiltonian(eps_imp, U, eps_chain, v_chain, n_up, n_dn):
    n = len(eps_chain) + 1
    states = sector_states(n, n_up, n_dn)
    pos = {s: k for k, s in enumerate(states)}
    onsite = np.concatenate([[eps_imp], eps_chain])
    H = np.zeros((len(states), len(states)))
    for k, s in enumerate(states):
        e = 0.0
        for i in range(n):
            e += onsite[i] * (((s >> (2 * i)) & 1) + ((s >> (2 * i + 1)) & 1))
        if (s & 1) and (s & 2):
            e += U
        H[k, k] += e
        for b, vb in enumerate(v_chain):  # bond (b, b+1)
            for sp in (0, 1):
                a, c = 2 * b + sp, 2 * (b + 1) + sp
                for i, j in ((a, c), (c, a)):
                    r = _cdag_c(s, i, j)
                    if r is not None:
                        H[pos[r[1]], k] += vb * r[0]
    return states, H


def _apply_c(states_from, psi, states_to, orb, dagger):
    pos = {s: k for k, s in enumerate(states_to)}
    out = np.zeros(len(states_to))
    for k, s in enumerate(states_from):
        occ = (s >> orb) & 1
        if dagger == bool(occ):
            continue
        sign = -1 if _popcount(s & ((1 << orb) - 1)) % 2 else 1
        t = s ^ (1 << orb)
        out[pos[t]] += sign * psi[k]
    return out


def aim_ground_state(eps_imp, U, eps_chain, v_chain, n_up, n_dn):
    states, H = aim_sector_hamiltonian(eps_imp, U, eps_chain, v_chain, n_up, n_dn)
    w, V = np.linalg.eigh(H)
    return w[0], V[:, 0], states


def aim_green(eps_imp, U, eps_chain, v_chain, n_up, n_dn, z):
    """G_up(z) of the impurity from full diagonalisation of the N+-1 sectors."""
    e0, psi, st = aim_ground_state(eps_imp, U, eps_chain, v_chain, n_up, n_dn)
    z = np.asarray(z, dtype=np.complex128)
    g = np.zeros(z.shape, dtype=np.complex128)
    for dagger, sec, sgn in ((True, (n_up + 1, n_dn), 1), (False, (n_up - 1, n_dn), -1)):
        if sec[0] < 0 or sec[0] > len(eps_chain) + 1:
            continue
        st2, H2 = aim_sector_hamiltonian(eps_imp, U, eps_chain, v_chain, *sec)
        phi = _apply_c(st, psi, st2, 0, dagger)
        w, V = np.linalg.eigh(H2)
        amp = (V.T @ phi) ** 2
        for a, e in zip(amp, w):
            g += a / (z - sgn * (e - e0))
    return e0, g


def aim_green_time(eps_imp, U, eps_chain, v_chain, n_up, n_dn, t):
    """Retarded G_up(t) = -i <{c_0up(t), c+_0up}> from the same full
    diagonalisation of the N+-1 sectors as aim_green:
    -i sum_k a_k exp(-i p_k t) over the poles p_k and weights a_k of G(z)
    (the real-time object of tk-dmft/src/spectral/tdvp.rs:55-108)."""
    e0, psi, st = aim_ground_state(eps_imp, U, eps_chain, v_chain, n_up, n_dn)
    t = np.asarray(t, dtype=np.float64)
    g = np.zeros(t.shape, dtype=np.complex128)
    for dagger, sec, sgn in ((True, (n_up + 1, n_dn), 1), (False, (n_up - 1, n_dn), -1)):
        if sec[0] < 0 or sec[0] > len(eps_chain) + 1:
            continue
        st2, H2 = aim_sector_hamiltonian(eps_imp, U, eps_chain, v_chain, *sec)
        phi = _apply_c(st, psi, st2, 0, dagger)
        w, V = np.linalg.eigh(H2)
        amp = (V.T @ phi) ** 2
        p = sgn * (w - e0)
        g += -1j * (np.exp(-1j * np.multiply.outer(t, p)) @ amp)
    return e0, g


# ---------------------------------------------------------------- MLP
def mlp_unpack(p, dims):
    out, o = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = p[o:o + a * b].reshape(a, b)
        o += a * b
        bb = p[o:o + b]
        o += b
        out.append((W, bb))
    return out


def mlp_forward(p, dims, x):
    p = np.asarray(p, dtype=np.float64)
    a = np.asarray(x, dtype=np.float64)
    layers = mlp_unpack(p, dims)
    for l, (W, b) in enumerate(layers):
        z = a @ W + b
        a = np.tanh(z) if l + 1 < len(layers) else z
    return a


def mlp_loss_grad(p, dims, x, t):
    p = np.asarray(p, dtype=np.float64)
    layers = mlp_unpack(p, dims)
    acts = [np.asarray(x, dtype=np.float64)]
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ W + b
        acts.append(np.tanh(z) if l + 1 < len(layers) else z)
    y = acts[-1]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    G = 2.0 * d / d.size
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gW = acts[l].T @ G
        gb = G.sum(0)
        grads.append((gW, gb))
        if l > 0:
            G = (G @ W.T) * (1 - acts[l] ** 2)
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb]
    return loss, np.concatenate(flat)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
