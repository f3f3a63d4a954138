"""DMFT pieces on the GPU against the oracle (dmft_oracle.py).

Bars: hybridization functions within 1e-13 relative of the oracle; the host
discretiser and Broyden mixer equal to the oracle's restatement within 1e-14;
the device least-squares fit matches an exactly representable bath to 1e-10 in relative cost; the DMRG
impurity ground state within 1e-10 of exact diagonalisation and G(i w_n)
within 1e-9; the FP32 network against a float64 restatement (forward 1e-5,
gradients 1e-4 relative to the largest entry) and one Adam step.
"""
import math

import numpy as np
import pytest

import dmft_oracle as O
import tk_dmft as D

pytestmark = pytest.mark.gpu


def test_hybridization_reference_value(ctx):
    print("case dmft hybridization known value", flush=True)
    bath = D.BathParameters([-1.0, 1.0], [0.5, 0.5])
    d = bath.hybridization_function(ctx, [0.0], 0.1)
    expected = 0.25 / complex(1.0, 0.1) + 0.25 / complex(-1.0, 0.1)
    assert abs(d[0] - expected) < 1e-12
    u = D.BathParameters.uniform(4, 10.0, 1.0)
    om = np.arange(-50, 51) * 0.2
    assert u.hybridization_distance(ctx, u, om, 0.05) < 1e-12
    m = D.BathParameters([-1.0, 1.0], [1.0, 1.0]).linear_mix(D.BathParameters([-2.0, 2.0], [0.5, 0.5]), 0.3)
    assert np.allclose(m.epsilon, [-1.3, 1.3], atol=1e-12) and abs(m.v[0] - 0.85) < 1e-12


@pytest.mark.parametrize("geometry", [D.STAR, D.CHAIN])
def test_hybridization_batched_vs_oracle(ctx, geometry):
    print(f"case dmft hybridization batched geometry={geometry}", flush=True)
    rng = np.random.default_rng(11)
    P, nb = 32, 7
    params = np.concatenate([rng.uniform(-2, 2, (P, nb)), rng.uniform(0.1, 0.8, (P, nb))], 1)
    z = np.concatenate([1j * D.matsubara(50.0, 256), np.linspace(-3, 3, 64) + 0.05j])
    got = D.hybridization(ctx, geometry, params, z)
    f = O.hyb_star if geometry == D.STAR else O.hyb_chain
    for p in range(P):
        ref = f(params[p, :nb], params[p, nb:], z)
        assert np.max(np.abs(got[p] - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_discretiser_and_broyden_vs_oracle(ctx):
    print("case dmft discretiser broyden", flush=True)
    omega = np.array([-5.0 + 10.0 * i / 199.0 for i in range(200)])
    im = np.where(np.abs(omega) < 2, (2 / math.pi) * np.sqrt(np.clip(1 - (omega / 2) ** 2, 0, None)), 0.0)
    bath = D.lanczos_tridiagonalize(im, omega, 8)
    dw = omega[1] - omega[0]
    ref_e, ref_v = O.lanczos_points(omega, im * dw / math.pi, 8)
    assert np.allclose(bath.epsilon, ref_e, atol=1e-14, rtol=1e-14)
    assert np.allclose(bath.v, ref_v, atol=1e-14, rtol=1e-14)
    assert np.all(np.abs(bath.epsilon) <= 6.0) and bath.v[0] > 0
    # star -> chain keeps Delta
    rng = np.random.default_rng(3)
    star = D.BathParameters(np.sort(rng.uniform(-2, 2, 7)), rng.uniform(0.1, 0.6, 7))
    ce, cv = D.star_to_chain(star)
    z = 1j * D.matsubara(50.0, 64)
    a = D.hybridization(ctx, D.CHAIN, np.concatenate([ce, cv]), z)[0]
    b = D.hybridization(ctx, D.STAR, star.params(), z)[0]
    assert np.max(np.abs(a - b)) < 1e-12
    # Broyden: same sequence as the restatement
    bg, bo = D.BroydenState(3), O.Broyden(3)
    x = np.array([0.3, -0.2, 0.5, 0.1])
    for it in range(6):
        f = np.tanh(x) + 0.1 * np.sin(3 * x + it)
        xg, xo = bg.update(x, f, 0.5), bo.update(x, f, 0.5)
        assert np.allclose(xg, xo, atol=1e-14, rtol=1e-13)
        x = xg


def test_bath_fit_recovers_bath(ctx):
    print("case dmft bath fit", flush=True)
    rng = np.random.default_rng(21)
    wn = D.matsubara(50.0, 256)
    P, nb = 8, 7
    truth = np.concatenate([np.sort(rng.uniform(-2, 2, (P, nb)), 1), rng.uniform(0.2, 0.6, (P, nb))], 1)
    delta = D.hybridization(ctx, D.STAR, truth, 1j * wn)
    start = truth + rng.normal(0, 0.05, truth.shape)
    fit, cost = D.bath_fit(ctx, wn, delta, start, max_iter=1000)
    for p in range(P):
        d_fit = O.hyb_star(fit[p, :nb], fit[p, nb:], 1j * wn)
        ref = np.sum(np.abs(d_fit - delta[p]) ** 2)
        assert cost[p] == pytest.approx(ref, rel=1e-6, abs=1e-24 * np.sum(np.abs(delta[p]) ** 2))
        # Matsubara-axis fits are ill-conditioned (J^T J solved by Cholesky):
        # the data are matched to ~1e-6 relative
        assert cost[p] < 1e-10 * np.sum(np.abs(delta[p]) ** 2)


def test_aim_ground_state_and_green_vs_ed(ctx):
    print("case dmft impurity n_bath=3", flush=True)
    rng = np.random.default_rng(4)
    U, nb = 4.0, 3
    eps = rng.uniform(-1, 1, nb)
    v = rng.uniform(0.3, 0.9, nb)
    solver = D.ImpuritySolver(ctx, nb, U, bond_dim=64, sweeps=8)
    wn = D.matsubara(50.0, 64)
    e0, g = solver.green(eps, v, wn)
    e_ref, g_ref = O.aim_green(-U / 2, U, eps, v, 2, 2, 1j * wn)
    assert abs(e0 - e_ref) < 1e-10
    assert np.max(np.abs(g - g_ref)) < 1e-9


def test_aim_c5_ground_state_vs_ed(ctx):
    """The 8-site impurity (16 spin-orbitals) at D = 256: DMRG is exact."""
    print("case dmft impurity n_bath=7 D=256", flush=True)
    rng = np.random.default_rng(8)
    U, nb = 4.0, 7
    eps = np.sort(rng.uniform(-1.5, 1.5, nb))
    v = rng.uniform(0.3, 0.8, nb)
    solver = D.ImpuritySolver(ctx, nb, U, bond_dim=256, sweeps=6)
    e0, psi, (nu, nd) = solver.ground_state(eps, v)
    # exact energy of the N=8, Sz=0 sector (4,900 states) by the oracle's ED
    basis, rowptr, col, val = D.sector_csr(-U / 2, U, eps, v, nu, nd)
    assert basis.size == 4900
    st, H = O.aim_sector_hamiltonian(-U / 2, U, eps, v, nu, nd)
    e_ref = np.linalg.eigvalsh(H)[0]
    assert abs(e0 - e_ref) < 1e-10
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-10


def test_aim_c5_green_time_tdvp_vs_ed(ctx):
    """Config-5 G(t): the 8-site impurity (7 bath sites, D = 256) by real-time
    two-site TDVP to t = 20 in steps of 0.05 (400 steps per part), particle
    plus hole, against exact diagonalisation (1e-10; D = 256 holds the N+-1
    states exactly, so only the Krylov exponentials err)."""
    print("case dmft impurity G(t) n_bath=7 t=20", flush=True)
    rng = np.random.default_rng(8)
    U, nb = 4.0, 7
    eps = np.sort(rng.uniform(-1.5, 1.5, nb))
    v = rng.uniform(0.3, 0.8, nb)
    solver = D.ImpuritySolver(ctx, nb, U, bond_dim=256, sweeps=6)
    t, g, info = solver.green_time(eps, v, dt=0.05, t_max=20.0)
    assert len(t) == 401 and abs(t[-1] - 20.0) < 1e-12
    e_ref, g_ref = O.aim_green_time(-U / 2, U, eps, v, 4, 4, t)
    assert abs(info["e0"] - e_ref) < 1e-10
    assert abs(info["weight_particle"] + info["weight_hole"] - 1.0) < 1e-12
    assert np.max(np.abs(g - g_ref)) < 1e-10


def test_green_time_many_concurrent(ctx):
    """Independent impurity problems solved by four threads on their own
    streams give the sequential results bit for bit, and ED's G(t)."""
    print("case dmft concurrent G(t)", flush=True)
    rng = np.random.default_rng(12)
    U, nb = 4.0, 3
    baths = [(np.sort(rng.uniform(-1.5, 1.5, nb)), rng.uniform(0.3, 0.8, nb)) for _ in range(6)]
    seq = D.green_time_many(baths, nb, U, workers=1, bond_dim=64, t_max=2.0)
    par = D.green_time_many(baths, nb, U, workers=4, bond_dim=64, t_max=2.0)
    for (eps, v), (t, g, _), (t2, g2, _) in zip(baths, seq, par):
        assert np.array_equal(g, g2)
        _, g_ref = O.aim_green_time(-U / 2, U, eps, v, 2, 2, t)
        assert np.max(np.abs(g - g_ref)) < 1e-10


def test_mlp_forward_gradient_adam_vs_oracle(ctx):
    print("case dmft mlp", flush=True)
    dims = [512, 256, 256, 14]
    net = D.Mlp(ctx, dims, seed=3)
    assert net.n_params == 512 * 256 + 256 + 256 * 256 + 256 + 256 * 14 + 14
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(9)
    x = rng.normal(size=(16, 512)).astype(np.float32)
    t = rng.normal(size=(16, 14)).astype(np.float32)
    y = net.forward(x)
    y_ref = O.mlp_forward(p, dims, x.astype(np.float64))
    assert np.max(np.abs(y - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref = O.mlp_loss_grad(p, dims, x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * np.max(np.abs(g_ref))
    # one Adam step from zero moments, then a second with a new gradient (the
    # oracle's Adam is pinned by the two-step closed form in test_oracle_dmft)
    net.adam(1e-3)
    p1 = net.params().astype(np.float64)
    p_ref, m_ref, v_ref = O.adam(p, g.astype(np.float64), np.zeros_like(p), np.zeros_like(p), 1, 1e-3)
    assert np.max(np.abs(p1 - p_ref)) <= 1e-6
    x2 = rng.normal(size=(16, 512)).astype(np.float32)
    _, g2 = net.loss_grad(x2, t)
    net.adam(1e-3)
    p2 = net.params().astype(np.float64)
    p_ref2, _, _ = O.adam(p1, g2.astype(np.float64), m_ref, v_ref, 2, 1e-3)
    assert np.max(np.abs(p2 - p_ref2)) <= 1e-6
    # training lowers the loss
    hist = net.train(x, t, epochs=60, lr=1e-3)
    assert hist[-1] < 0.2 * hist[0]


def test_dmft_loop_and_learned_proposal(ctx):
    """Config-5 impurity (7 bath sites, D = 256, U around 4, beta = 50, 256
    Matsubara points): the self-consistency loop converges, and the network
    trained on the solver's own (Delta, converged bath) pairs from 16 other
    U values reaches the same baths in fewer impurity solves over five
    held-out U values (a rejected proposal costs one judging solve): at
    least a quarter fewer in total (measured: about half)."""
    print("case dmft loop + learned proposal", flush=True)
    runs = []
    for k in range(16):
        U = 3.0625 + 0.125 * k
        loop = D.DMFTLoop(ctx, n_bath=7, U=U, tol=1e-5, max_iterations=60)
        chain, n, ok = loop.solve()
        assert ok, f"U={U} did not converge"
        runs.append((loop.history, chain))
    net, proposal, loss = D.train_bath_proposal(ctx, runs)
    assert loss[-1] < 1e-4 * loss[0]
    tot_plain = tot_learned = 0
    for U in (3.9, 4.0, 4.1, 4.25, 4.4):
        plain = D.DMFTLoop(ctx, n_bath=7, U=U, tol=1e-5, max_iterations=60)
        ca, na, oka = plain.solve()
        learned = D.DMFTLoop(ctx, n_bath=7, U=U, tol=1e-5, max_iterations=60)
        cb, nb, okb = learned.solve(proposal=proposal)
        print(f"U={U}: plain {na} solves, learned {nb} solves", flush=True)
        assert oka and okb
        assert np.max(np.abs(np.concatenate(ca) - np.concatenate(cb))) < 1e-3
        tot_plain += na
        tot_learned += nb
    print(f"total: plain {tot_plain}, learned {tot_learned}", flush=True)
    assert tot_learned < 0.75 * tot_plain


def test_dmft_checkpoint_written_and_resumed(ctx, tmp_path):
    """A loop with a checkpoint path writes the reference's JSON after every
    iteration; resume_from_checkpoint returns that bath and rejects a bath
    of another size (loop/mod.rs:282-293)."""
    print("case dmft checkpoint", flush=True)
    path = tmp_path / "dmft.json"
    loop = D.DMFTLoop(ctx, n_bath=3, U=4.0, bond_dim=64, max_iterations=2, tol=1e-12, checkpoint_path=path)
    chain, n, ok = loop.solve()
    cp = D.DMFTCheckpoint.read_from_file(path)
    assert cp.iteration == 1 and not cp.converged
    assert len(cp.bath_epsilon) == 3 and len(cp.bath_v) == 3
    eps, v = loop.resume_from_checkpoint(cp)
    assert eps.shape == (3,) and np.all(np.isfinite(v))
    bad = D.DMFTCheckpoint(0, False, "{}", [0.0] * 4, [0.5] * 4)
    with pytest.raises(ValueError):
        loop.resume_from_checkpoint(bad)
