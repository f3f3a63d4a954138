"""CPU restatement of the reference's real-valued Krylov exponential.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

exp_krylov        tk-dmrg/src/tdvp/mod.rs:197-272 (exp_krylov_f64)
mat_exp_times_vec tk-dmrg/src/tdvp/mod.rs:278-382

Pinned by tests/test_oracle_tdvp.py: the reference's own three tests
(tdvp/mod.rs:410-484: identity operator, diag(1,2,3), zero time), and beyond
them scipy.linalg.expm (an independent Pade implementation) on random
symmetric matrices with a full Krylov space, the happy-breakdown exit, the
scaling-and-squaring branch (large ||alpha A||) and the semigroup law
exp(aA) exp(bA) v = exp((a+b)A) v.
"""
from __future__ import annotations

import math

import numpy as np


def mat_exp_times_vec(a, beta):
    """exp(a) (beta e_1) for a small n x n matrix a (tdvp/mod.rs:278-382).

    Scaling exponent s = max(0, ceil(log2(sum|a_ij| / n / 0.5))); with s = 0 a
    Taylor series on the vector (at most 20 terms, stop when the term's
    2-norm falls below 1e-16 |beta|); with s > 0 the full matrix Taylor series
    of a 2^-s (stop when the term's Frobenius norm falls below 1e-16), squared
    s times, first column times beta."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    if n == 0:
        return np.zeros(0)
    norm = float(np.abs(a).sum()) / n
    s = max(0, math.ceil(math.log2(norm / 0.5))) if norm > 0 else 0
    scale = 2.0 ** (-s)
    if s == 0:
        result = np.zeros(n)
        result[0] = beta
        term = result.copy()
        for k in range(1, 21):
            term = (a * scale) @ term / k
            result = result + term
            if np.linalg.norm(term) < 1e-16 * abs(beta):
                break
        return result
    e = np.eye(n)
    t = np.eye(n)
    for k in range(1, 21):
        t = (a * scale) @ t / k
        e = e + t
        if np.linalg.norm(t) < 1e-16:
            break
    for _ in range(s):
        e = e @ e
    return e[:, 0] * beta


def exp_krylov(matvec, v, alpha, krylov_dim, tol):
    """exp(alpha A) v by Arnoldi with modified Gram-Schmidt (tdvp/mod.rs:197-272).

    m = min(krylov_dim, n) Arnoldi steps; stop early when h_{j+1,j} < tol
    (happy breakdown); a zero vector (||v|| < 1e-30) maps to zero.  Returns
    (result, krylov dimension used)."""
    v = np.asarray(v, dtype=np.float64)
    n = v.shape[0]
    m = min(krylov_dim, n)
    if m == 0 or n == 0:
        return v.copy(), 0
    beta = float(np.sqrt(np.sum(v * v)))
    if beta < 1e-30:
        return np.zeros(n), 0
    basis = [v / beta]
    h = np.zeros((m + 1, m))
    used = m
    for j in range(m):
        w = np.asarray(matvec(basis[j]), dtype=np.float64).copy()
        for i in range(j + 1):
            d = float(basis[i] @ w)
            h[i, j] = d
            w = w - d * basis[i]
        hn = float(np.sqrt(np.sum(w * w)))
        h[j + 1, j] = hn
        if hn < tol:
            used = j + 1
            break
        basis.append(w / hn)
    c = mat_exp_times_vec(alpha * h[:used, :used], beta)
    out = np.zeros(n)
    for i in range(used):
        out = out + c[i] * basis[i]
    return out, used
