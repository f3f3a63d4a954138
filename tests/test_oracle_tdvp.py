"""Pins the Krylov-exponential oracle (oracle/tdvp_oracle.py).

The reference's own tests (tk-dmrg/src/tdvp/mod.rs:410-484) first, then
scipy.linalg.expm (Pade, independent of the Taylor/squaring restatement) where
the Krylov space is complete, the breakdown exit and the semigroup law."""
import os
import sys

import numpy as np
import pytest
import scipy.linalg

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import tdvp_oracle as to  # noqa: E402


def test_reference_identity_operator():
    # tdvp/mod.rs:410-431: exp(0.5 I) e_1 = e^0.5 e_1
    r, m = to.exp_krylov(lambda x: x.copy(), np.array([1.0, 0, 0, 0, 0]), 0.5, 10, 1e-12)
    assert abs(r[0] - np.exp(0.5)) < 1e-8
    assert np.all(np.abs(r[1:]) < 1e-10)
    assert m == 1  # A e_1 = e_1: breakdown after one step


def test_reference_diagonal_matrix():
    # tdvp/mod.rs:433-458: exp(-0.1 diag(1,2,3)) [1,1,1]
    d = np.array([1.0, 2.0, 3.0])
    r, _ = to.exp_krylov(lambda x: d * x, np.ones(3), -0.1, 10, 1e-12)
    assert np.max(np.abs(r - np.exp(-0.1 * d))) < 1e-6
    # the restatement is far tighter than the reference's bound
    assert np.max(np.abs(r - np.exp(-0.1 * d))) < 1e-14


def test_reference_zero_time():
    # tdvp/mod.rs:460-483: exp(0 A) v = v for the 1-D Laplacian
    a = 2 * np.eye(4) - np.eye(4, k=1) - np.eye(4, k=-1)
    v = np.array([1.0, 2, 3, 4])
    r, _ = to.exp_krylov(lambda x: a @ x, v, 0.0, 10, 1e-12)
    assert np.max(np.abs(r - v)) < 1e-10


@pytest.mark.parametrize("n,alpha", [(6, -0.3), (12, 0.2), (20, -1.0), (30, -0.05)])
def test_full_krylov_matches_expm(n, alpha):
    rng = np.random.default_rng(n)
    g = rng.standard_normal((n, n))
    a = (g + g.T) / 2
    v = rng.standard_normal(n)
    r, m = to.exp_krylov(lambda x: a @ x, v, alpha, n, 1e-14)
    ref = scipy.linalg.expm(alpha * a) @ v
    assert np.linalg.norm(r - ref) < 1e-11 * np.linalg.norm(ref)


def test_scaling_and_squaring_branch():
    rng = np.random.default_rng(3)
    g = rng.standard_normal((8, 8))
    a = 3 * (g + g.T)
    for beta in (1.0, 2.5):
        c = to.mat_exp_times_vec(-0.7 * a, beta)  # ||a|| large: s > 0
        ref = scipy.linalg.expm(-0.7 * a)[:, 0] * beta
        assert np.linalg.norm(c - ref) < 1e-10 * np.linalg.norm(ref)
    small = 0.01 * a  # s = 0: vector Taylor series
    c = to.mat_exp_times_vec(small, 1.0)
    assert np.linalg.norm(c - scipy.linalg.expm(small)[:, 0]) < 1e-14


def test_semigroup_and_zero_vector():
    rng = np.random.default_rng(5)
    g = rng.standard_normal((16, 16))
    a = g + g.T
    v = rng.standard_normal(16)
    mv = lambda x: a @ x  # noqa: E731
    r1, _ = to.exp_krylov(mv, v, -0.1, 16, 1e-14)
    r2, _ = to.exp_krylov(mv, r1, -0.15, 16, 1e-14)
    r3, _ = to.exp_krylov(mv, v, -0.25, 16, 1e-14)
    assert np.linalg.norm(r2 - r3) < 1e-11 * np.linalg.norm(r3)
    z, m = to.exp_krylov(mv, np.zeros(16), -0.1, 16, 1e-14)
    assert m == 0 and not np.any(z)
