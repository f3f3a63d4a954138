"""Pins of the DMFT oracle against the reference's known answers and closed forms."""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import dmft_oracle as O  # noqa: E402


def test_uniform_bath_reference():
    # bath.rs:170-178 and :181-186
    e, v = O.uniform(6, 10.0, 1.0)
    assert e[0] == pytest.approx(-5.0, abs=1e-12) and e[5] == pytest.approx(5.0, abs=1e-12)
    assert np.all(np.abs(v - 1.0) < 1e-12)
    e1, _ = O.uniform(1, 10.0, 1.0)
    assert e1[0] == 0.0


def test_hybridization_known_value():
    # bath.rs:188-204: eps (-1, 1), v (0.5, 0.5), Delta(0 + 0.1i)
    d = O.hyb_star([-1.0, 1.0], [0.5, 0.5], np.array([0.1j]))
    expected = 0.25 / complex(1.0, 0.1) + 0.25 / complex(-1.0, 0.1)
    assert abs(d[0] - expected) < 1e-12


def test_linear_mix_reference():
    # bath.rs:206-223
    e, v = O.linear_mix(([-1.0, 1.0], [1.0, 1.0]), ([-2.0, 2.0], [0.5, 0.5]), 0.3)
    assert e[0] == pytest.approx(-1.3, abs=1e-12) and e[1] == pytest.approx(1.3, abs=1e-12)
    assert v[0] == pytest.approx(0.85, abs=1e-12)


def test_broyden_reference():
    # mixing.rs:181-229
    b = O.Broyden(5)
    x, f = np.array([1.0, 2.0, 3.0]), np.array([1.5, 2.5, 3.5])
    assert np.allclose(b.update(x, f, 0.3), 0.7 * x + 0.3 * f, atol=1e-12)
    b = O.Broyden(5)
    x1 = np.array([1.0, 2.0])
    x2 = b.update(x1, np.array([1.5, 2.5]), 0.5)
    f2 = x2 + 0.1
    x3 = b.update(x2, f2, 0.5)
    assert np.any(np.abs(x3 - (0.5 * x2 + 0.5 * f2)) > 1e-10)
    b = O.Broyden(2)
    x = np.array([1.0, 2.0])
    for _ in range(3):
        x = b.update(x, x + 0.1, 0.5)
    assert b.n_updates <= 2


def _semicircle_grid(n_omega):
    omega = np.array([-5.0 + 10.0 * i / (n_omega - 1.0) for i in range(n_omega)])
    half = 2.0
    w = np.where(np.abs(omega) < half, (2 / math.pi) * np.sqrt(np.clip(1 - (omega / half) ** 2, 0, None)), 0.0)
    return omega, w


def test_discretiser_reference():
    # discretize.rs:310-375: 4 and 8 bath sites of a semicircle; energies in
    # range, V_1 = sqrt(W) > 0
    for n_omega, n_bath in ((100, 4), (200, 8)):
        omega, im = _semicircle_grid(n_omega)
        dw = omega[1] - omega[0]
        eps, v = O.lanczos_points(omega, im * dw / math.pi, n_bath)
        assert eps.size == n_bath and v.size == n_bath
        assert np.all(np.abs(eps) <= 6.0)
        assert v[0] == pytest.approx(math.sqrt(np.sum(im * dw / math.pi)))
        assert v[0] > 0


def test_discretiser_exact_recovery():
    # a discrete star bath mapped to a chain keeps Delta(z) exactly
    rng = np.random.default_rng(3)
    eps = np.sort(rng.uniform(-2, 2, 7))
    v = rng.uniform(0.1, 0.6, 7)
    ce, cv = O.lanczos_points(eps, v ** 2, 7)
    z = 1j * (2 * np.arange(64) + 1) * math.pi / 50 + 0.3
    assert np.max(np.abs(O.hyb_chain(ce, cv, z) - O.hyb_star(eps, v, z))) < 1e-12
    # one site: v^2 / (z - eps)
    assert abs(O.hyb_chain([0.4], [0.7], np.array([1j]))[0] - 0.49 / (1j - 0.4)) < 1e-15


def test_ed_atomic_limit():
    U = 4.0
    z = 1j * np.array([0.1, 1.0, 3.0])
    e0, g = O.aim_green(-U / 2, U, [], [], 1, 0, z)
    # one spin-up electron on the impurity: only its removal, at -U/2
    assert e0 == pytest.approx(-U / 2)
    g_up = g
    g_expected = 1 / (z + U / 2)
    # the spin-down state: only the addition of an up electron, at +U/2
    _, gdn = O.aim_green(-U / 2, U, [], [], 0, 1, z)
    g_avg = 0.5 * (g_up + gdn)
    assert np.max(np.abs(g_avg - 0.5 * (1 / (z - U / 2) + 1 / (z + U / 2)))) < 1e-12
    assert np.max(np.abs(g_up - g_expected)) < 1e-12


def test_ed_free_fermions():
    rng = np.random.default_rng(5)
    eps_imp, eps, v = 0.2, rng.uniform(-1, 1, 3), rng.uniform(0.3, 0.8, 3)
    # one-particle chain matrix
    h = np.diag(np.concatenate([[eps_imp], eps]))
    for k in range(3):
        h[k, k + 1] = h[k + 1, k] = v[k]
    lev = np.linalg.eigvalsh(h)
    nocc = 2
    e0, g = O.aim_green(eps_imp, 0.0, eps, v, nocc, nocc, 1j * np.array([0.3, 1.1, 4.0]))
    assert e0 == pytest.approx(2 * lev[:nocc].sum(), abs=1e-12)
    z = 1j * np.array([0.3, 1.1, 4.0])
    g_free = 1 / (z - eps_imp - O.hyb_chain(eps, v, z))
    assert np.max(np.abs(g - g_free)) < 1e-12


def test_ed_green_time_free_fermions_and_limits():
    """U = 0: G^R_00(t) = -i [exp(-i h t)]_00 of the one-particle chain matrix
    (independent of the many-body ED); the atomic limit -i (e^{-i U t / 2}
    + e^{i U t / 2}) / 2; G^R(0) = -i (the anticommutator) for U = 4."""
    rng = np.random.default_rng(6)
    eps_imp, eps, v = 0.1, rng.uniform(-1, 1, 3), rng.uniform(0.3, 0.8, 3)
    h = np.diag(np.concatenate([[eps_imp], eps]))
    for k in range(3):
        h[k, k + 1] = h[k + 1, k] = v[k]
    t = np.linspace(0, 7, 29)
    _, g = O.aim_green_time(eps_imp, 0.0, eps, v, 2, 2, t)
    w, q = np.linalg.eigh(h)
    free = -1j * (np.exp(-1j * np.multiply.outer(t, w)) @ (q[0] ** 2))
    assert np.max(np.abs(g - free)) < 1e-12
    U = 4.0
    _, gu = O.aim_green_time(-U / 2, U, [], [], 1, 0, t)
    _, gd = O.aim_green_time(-U / 2, U, [], [], 0, 1, t)
    assert np.max(np.abs(0.5 * (gu + gd) + 1j * np.cos(U * t / 2))) < 1e-12
    _, g4 = O.aim_green_time(-U / 2, U, eps, v, 2, 2, [0.0])
    assert abs(g4[0] + 1j) < 1e-12


def test_mlp_gradient_finite_difference():
    dims = [6, 5, 4, 3]
    rng = np.random.default_rng(7)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = rng.normal(0, 0.4, n)
    x = rng.normal(size=(5, 6))
    t = rng.normal(size=(5, 3))
    loss, g = O.mlp_loss_grad(p, dims, x, t)
    assert loss == pytest.approx(np.mean((O.mlp_forward(p, dims, x) - t) ** 2))
    h = 1e-6
    for i in rng.choice(n, 25, replace=False):
        e = np.zeros(n)
        e[i] = h
        fd = (O.mlp_loss_grad(p + e, dims, x, t)[0] - O.mlp_loss_grad(p - e, dims, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-8 * max(1.0, abs(fd))


def test_adam_two_steps_closed_form():
    """The oracle's Adam (Kingma & Ba, bias-corrected) against the two-step
    closed form: from zero moments with gradients g1 then g2,
      step 1: m^ = g1, v^ = g1^2, p1 = p0 - lr g1 / (|g1| + eps)
      step 2: m^ = (b1 g1 + g2) / (1 + b1), v^ = (b2 g1^2 + g2^2) / (1 + b2),
              p2 = p1 - lr m^ / (sqrt(v^) + eps)
    (the bias corrections 1 - b^t with the moments' geometric sums); a zero
    gradient leaves its parameter, eps dominates a tiny one."""
    lr, b1, b2, eps = 0.1, 0.9, 0.999, 1e-8
    p0 = np.array([1.0, -2.0, 0.5, 3.0])
    g1 = np.array([0.5, -0.25, 0.0, 1e-12])
    g2 = np.array([-0.25, 2.0, 0.0, 1e-12])
    p, m, v = O.adam(p0, g1, np.zeros(4), np.zeros(4), 1, lr, b1, b2, eps)
    want1 = p0 - lr * g1 / (np.abs(g1) + eps)
    assert np.abs(p - want1).max() < 1e-15
    assert abs(want1[0] - (1.0 - 0.1 * 0.5 / (0.5 + 1e-8))) < 1e-16 and want1[2] == 0.5
    p, m, v = O.adam(p, g2, m, v, 2, lr, b1, b2, eps)
    mh = (b1 * g1 + g2) / (1 + b1)
    vh = (b2 * g1 ** 2 + g2 ** 2) / (1 + b2)
    want2 = want1 - lr * mh / (np.sqrt(vh) + eps)
    assert np.abs(p - want2).max() < 1e-14
    assert p[2] == 0.5
    # a gradient far below eps moves its parameter by ~lr |g| / eps
    assert abs((p0[3] - p[3]) - lr * 1e-12 / (1e-12 + eps) * 2) < 1e-6 * lr
