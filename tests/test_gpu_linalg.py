"""The dense LinAlgBackend methods and the split's block SVD on the GPU,
through the C ABI.

Known answers of the reference (tk-dmrg/src/truncation/mod.rs:185-236,
tk-linalg/src/device/faer.rs:795-930, proptest_linalg.rs:88-129), LAPACK
(numpy) for spectra and subspaces, and the library's own selection rule
(pinned in test_truncation.py) for the kept counts of the block SVD.
"""
import threading
import time

import numpy as np
import pytest

import tk_hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be(ctx):
    return tk_hip.HipBackend(ctx)


def _orth(rng, n, k):
    q, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return q


def _sigma_ok(s, s_true, smax):
    """Accuracy of sigma through the two Gram levels: a value the first level
    resolves (sigma >= 3.2e-5 sigma_max) comes from an eigenvalue of X X^T
    accurate to ~eps sigma_max^2, so |d sigma| <= 1e-14 sigma_max^2 / sigma;
    below, the second level (X with the first level's vectors projected out)
    is accurate to ~eps sigma_max in sigma itself: |d sigma| <= 1e-13 sigma_max."""
    tol = np.where(s_true >= 3.2e-5 * smax, 1e-14 * smax ** 2 / s_true, 1e-13 * smax)
    return bool(np.all(np.abs(s - s_true) <= tol))


def test_truncate_svd_reference_cases(be):
    print("case truncate_svd known answers", flush=True)
    # truncation/mod.rs:196-212
    a = np.array([[3.0, 0.0], [0.0, 2.0], [0.0, 0.0]])
    u, s, vt, bd, nd, err = be.truncate_svd(a, 200, 1e-12, 1)
    assert (bd, nd) == (2, 0) and err < 1e-14
    assert abs(s[0] - 3) < 1e-10 and abs(s[1] - 2) < 1e-10
    assert u.shape == (3, 2) and vt.shape == (2, 2)
    assert np.abs(u @ np.diag(s) @ vt - a).max() < 1e-12
    # truncation/mod.rs:214-227
    a = np.diag([3.0, 2.0, 1.0])
    u, s, vt, bd, nd, err = be.truncate_svd(a, 2, 1e-12, 1)
    assert (bd, nd) == (2, 1) and abs(err - 1.0 / 14.0) < 1e-10
    # truncation/mod.rs:229-236
    a = np.diag([10.0, 1e-15])
    u, s, vt, bd, nd, err = be.truncate_svd(a, 200, 1e-10, 2)
    assert bd >= 2 and u.shape == (2, 2)
    assert np.abs(u.T @ u - np.eye(2)).max() < 1e-12 and np.abs(vt @ vt.T - np.eye(2)).max() < 1e-12


def test_svd_truncated_faer_cases(be):
    print("case svd_truncated faer known answers", flush=True)
    a = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])
    # faer.rs:795-826: full rank, reconstruction < 1e-10
    u, s, vt, te = be.svd_truncated(a, 2, 0.0)
    assert len(s) == 2 and s[0] >= s[1]
    assert np.abs(u @ np.diag(s) @ vt - a).max() < 1e-10
    # faer.rs:828-843: rank 1, truncation error > 0 (= sigma_2^2)
    u, s, vt, te = be.svd_truncated(a, 1, 0.0)
    sl = np.linalg.svd(a, compute_uv=False)
    assert len(s) == 1 and te > 0 and abs(te - sl[1] ** 2) < 1e-12 * sl[0] ** 2


def test_svd_round_trip_property(be):
    # proptest_linalg.rs:88-129: m, n in 2..16, relative Frobenius error < 1e-10
    print("case svd round trips", flush=True)
    rng = np.random.default_rng(0)
    for m in (2, 3, 7, 16):
        for n in (2, 5, 16):
            a = rng.uniform(-1, 1, (m, n))
            u, s, vt, te = be.svd_truncated(a, min(m, n), 0.0)
            assert np.linalg.norm(u @ np.diag(s) @ vt - a) / np.linalg.norm(a) < 1e-10


@pytest.mark.parametrize("m,n", [(60, 90), (90, 60)])
def test_svd_truncated_resolves_the_default_cutoff(be, m, n):
    """sigma spread over 1 .. 1e-11 (the reference keeps sigma >= 1e-12
    sigma_max by default, sweep/mod.rs:151): every value above the cutoff is
    kept (to the accuracy of _sigma_ok), none below is."""
    print(f"case svd graded spectrum {m}x{n}", flush=True)
    rng = np.random.default_rng(m * 1000 + n)
    k = min(m, n)
    sig = np.sort(10.0 ** rng.uniform(-14.5, 0, k))[::-1]
    sig = np.where(np.abs(np.log10(sig) + 12) < 0.5, sig * 10, sig)  # keep values away from the cut
    sig = np.sort(sig)[::-1]
    a = _orth(rng, m, k) @ np.diag(sig) @ _orth(rng, n, k).T
    u, s, vt, te = be.svd_truncated(a, k, 1e-12)
    want = sig[sig >= 1e-12 * sig[0]]
    assert len(s) == len(want)
    assert _sigma_ok(s, want, sig[0])
    assert np.abs(u.T @ u - np.eye(len(s))).max() < 1e-10
    assert np.linalg.norm(u @ np.diag(s) @ vt - a) <= 1e-11 * sig[0]


def test_svd_rejects_oversize(be):
    with pytest.raises(tk_hip.TkError) as e:
        be.svd_truncated(np.zeros((2049, 2049)), 10, 0.0)
    assert e.value.status == tk_hip.TK_ERR_INVALID_ARG


def test_block_svd_known_answers(be):
    """The split's block SVD: sigma (3, 1) and (2) cut to bond 2 keep one
    value of each block with error 1/14; a pair degenerate across blocks is
    never split."""
    print("case block svd known answers", flush=True)
    rng = np.random.default_rng(7)
    b0 = _orth(rng, 2, 2) @ np.diag([3.0, 1.0]) @ _orth(rng, 3, 2).T  # 2 x 3, sigma (3, 1)
    b1 = _orth(rng, 1, 1) @ np.diag([2.0]) @ _orth(rng, 4, 1).T       # 1 x 4, sigma (2)
    out, nk, nd, err = be.block_svd_truncate([b0, b1], 2, 1e-12)
    assert [len(o[1]) for o in out] == [1, 1] and (nk, nd) == (2, 1)
    assert abs(err - 1.0 / 14.0) < 1e-12
    assert abs(out[0][1][0] - 3) < 1e-12 and abs(out[1][1][0] - 2) < 1e-12
    b0 = _orth(rng, 3, 2) @ np.diag([1.0, 0.5]) @ _orth(rng, 3, 2).T
    b1 = _orth(rng, 2, 1) @ np.diag([0.5 - 1e-14]) @ _orth(rng, 2, 1).T
    out, nk, nd, err = be.block_svd_truncate([b0, b1], 2, 0.0, degeneracy_tol=1e-12)
    assert [len(o[1]) for o in out] == [1, 0]
    out, nk, nd, err = be.block_svd_truncate([b0, b1], 2, 0.0, degeneracy_tol=0.0)
    assert nk == 2


def test_block_svd_matches_lapack(be):
    print("case block svd vs LAPACK", flush=True)
    rng = np.random.default_rng(11)
    blocks, sig = [], []
    for (m, n) in [(40, 30), (25, 60), (70, 70), (5, 9), (1, 3)]:
        k = min(m, n)
        s = 10.0 ** rng.uniform(-13, 0, k)
        # the second Gram level resolves sigma to ~1e-13 sigma_max, so keep the
        # spectrum out of [10^-12.5, 10^-11.5] where the 1e-12 cut could flip
        s = np.sort(np.where(np.abs(np.log10(s) + 12) < 0.5, s * 10, s))[::-1]
        blocks.append(_orth(rng, m, k) @ np.diag(s) @ _orth(rng, n, k).T)
        sig.append(np.linalg.svd(blocks[-1], compute_uv=False))
    for max_bond in (10, 60, 500):
        out, nk, nd, err = be.block_svd_truncate(blocks, max_bond, 1e-12)
        keep, nk2, nd2, err2 = tk_hip.truncation_select(sig, max_bond, 1e-12, 1, 1e-12, 1e-26)
        assert [len(o[1]) for o in out] == keep and nk == nk2 and nd == nd2
        assert abs(err - err2) <= 1e-13
        for b, (u, s, c) in zip(blocks, out):
            k = len(s)
            if k == 0:
                continue
            ul, sl, _ = np.linalg.svd(b)
            assert _sigma_ok(s, sl[:k], max(x[0] for x in sig))
            assert np.abs(u.T @ u - np.eye(k)).max() < 1e-10
            assert np.abs(c - u.T @ b).max() <= 1e-12 * sl[0]
            # the kept subspace (where the k-th value is separated from the next)
            if k < len(sl) and sl[k - 1] - sl[k] > 1e-3 * sl[0]:
                pu = u @ u.T
                pl = ul[:, :k] @ ul[:, :k].T
                assert np.abs(pu - pl).max() < 1e-8


@pytest.mark.parametrize("n,mults", [(10, (4,)), (64, (7, 3)), (200, (7, 5, 2)), (700, (12, 4))])
def test_eigh_exactly_degenerate(be, n, mults):
    """Exactly repeated eigenvalues (their copies sit in separate split blocks
    of the tridiagonal form): every vector is an eigenvector and the set is
    orthonormal (LAPACK-level: residual <= 1e-12 |A|)."""
    print(f"case eigh degenerate n={n} {mults}", flush=True)
    rng = np.random.default_rng(n)
    lam = list(rng.uniform(-1, 1, n - sum(mults)))
    for j, m in enumerate(mults):
        lam += [0.25 * (j + 1) - 0.6] * m
    q = _orth(rng, n, n)
    a = (q * np.array(lam)) @ q.T
    a = (a + a.T) / 2
    w, v = be.eigh_lowest(a, n)
    assert np.abs(w - np.sort(lam)).max() < 1e-13 * n
    assert np.abs(a @ v - v * w).max() < 1e-12 * np.abs(a).max() * np.sqrt(n)
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("nb,m", [(8, 4), (33, 3), (5, 40)])
def test_eigh_block_diagonal_copies(be, nb, m):
    """m identical diagonal blocks, rows permuted: the tridiagonal form splits
    exactly (zero off-diagonals) into copies with bit-equal eigenvalues."""
    print(f"case eigh block copies {nb}x{m}", flush=True)
    rng = np.random.default_rng(nb * 100 + m)
    b = rng.standard_normal((nb, nb))
    b = b + b.T
    a = np.kron(np.eye(m), b)
    p = rng.permutation(nb * m)
    a = a[np.ix_(p, p)]
    n = nb * m
    w, v = be.eigh_lowest(a, n)
    assert np.abs(w - np.linalg.eigvalsh(a)).max() < 1e-13 * np.abs(a).max() * n
    assert np.abs(a @ v - v * w).max() < 1e-12 * np.abs(a).max() * np.sqrt(n)
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12


def test_block_svd_degenerate_reconstructs(be):
    """Blocks whose singular values repeat exactly (the bond blocks of a sum of
    identical operators): U has orthonormal columns and U (U^T X) = X."""
    print("case block svd degenerate", flush=True)
    rng = np.random.default_rng(5)
    blocks = []
    for m, n, s in [(10, 6, [1, .5, .5, .5, .5, .298]), (16, 9, [1] + [.5] * 7 + [.444]),
                    (40, 30, [2.0] * 10 + list(np.geomspace(1, 1e-6, 20)))]:
        k = len(s)
        blocks.append(_orth(rng, m, k) @ np.diag(s) @ _orth(rng, n, k).T)
    out, nk, nd, err = be.block_svd_truncate(blocks, 10 ** 6, 0.0, degeneracy_tol=0.0)
    for b, (u, s, c) in zip(blocks, out):
        assert np.abs(u.T @ u - np.eye(len(s))).max() < 1e-12
        # the third block spans sigma 2 .. 1e-6: its directions near the
        # first Gram level's line (3e-5 sigma_max) come from the second level,
        # which recovers the span to a few thousand eps |X| (measured 0.7 to
        # 1.6e-12 |X| across the summation orders of the reductions)
        assert np.abs(u @ c - b).max() < 4e-12 * np.abs(b).max()


def _graded(rng, k, n_big):
    """n_big values spread over 1 .. 3e-5 (above the first Gram level's line)
    and the rest over 2.5e-5 .. 3e-15, kept away from the 1e-12 cut."""
    big = 10.0 ** np.linspace(0, -4.4, n_big)
    small = 10.0 ** rng.uniform(-14.5, -4.6, k - n_big)
    small = np.where(np.abs(np.log10(small) + 12) < 0.5, small * 10, small)
    return np.sort(np.concatenate([big, small]))[::-1]


@pytest.mark.parametrize("flat", [False, True])
def test_block_svd_large_blocks_match_lapack(be, flat):
    """Blocks of 256 to 700 rows with graded spectra (few values above the
    first Gram level's line, 3e-5 sigma_max) or flat ones: LAPACK's kept
    counts, values to the accuracy of _sigma_ok, an orthonormal U with
    C = U^T X, the kept subspace, and the first level's directions to 5e-8."""
    print(f"case block svd large blocks flat={flat}", flush=True)
    rng = np.random.default_rng(31 + flat)
    blocks, sig = [], []
    for (m, n, nbig) in [(500, 600, 20), (300, 280, 35), (700, 700, 12), (256, 300, 50)]:
        k = min(m, n)
        s = 10.0 ** rng.uniform(-1, 0, k) if flat else _graded(rng, k, nbig)
        s = np.sort(s)[::-1]
        blocks.append(_orth(rng, m, k) @ np.diag(s) @ _orth(rng, n, k).T)
        sig.append(np.linalg.svd(blocks[-1], compute_uv=False))
    out, nk, nd, err = be.block_svd_truncate(blocks, 1500, 1e-12)
    keep, nk2, nd2, err2 = tk_hip.truncation_select(sig, 1500, 1e-12, 1, 1e-12, 1e-26)
    assert [len(o[1]) for o in out] == keep and nk == nk2 and nd == nd2
    smax = max(x[0] for x in sig)
    for b, (u, s, c), sl in zip(blocks, out, sig):
        k = len(s)
        assert _sigma_ok(s, sl[:k], smax)
        assert np.abs(u.T @ u - np.eye(k)).max() < 1e-10
        assert np.abs(c - u.T @ b).max() <= 1e-12 * sl[0]
        ul, _, _ = np.linalg.svd(b, full_matrices=False)
        if k < len(sl) and sl[k - 1] - sl[k] > 1e-3 * sl[0]:
            assert np.abs(u @ u.T - ul[:, :k] @ ul[:, :k].T).max() < 1e-8
        # the level-one directions (sigma >= 3.2e-5 sigma_max) to working precision
        k1 = int(np.sum(sl >= 3.2e-5 * smax))
        if 0 < k1 <= k:
            p1 = ul[:, :k1] @ ul[:, :k1].T
            assert np.abs(p1 @ u[:, :k1] - u[:, :k1]).max() < 5e-8


def test_eigh_lowest(be):
    print("case eigh_lowest", flush=True)
    # faer.rs:846-864
    a = np.array([[2.0, -1.0, 0.0], [-1.0, 2.0, -1.0], [0.0, -1.0, 2.0]])
    w, v = be.eigh_lowest(a, 2)
    assert len(w) == 2 and w[0] <= w[1] and abs(w[0] - (2 - np.sqrt(2))) < 1e-10
    assert np.abs(a @ v - v * w).max() < 1e-12 and v.shape == (3, 2)
    rng = np.random.default_rng(3)
    b = rng.standard_normal((150, 150))
    b = b + b.T
    w, v = be.eigh_lowest(b, 17)
    assert np.abs(w - np.linalg.eigvalsh(b)[:17]).max() < 1e-12 * np.abs(b).max() * 150
    assert np.abs(v.T @ v - np.eye(17)).max() < 1e-12
    with pytest.raises(tk_hip.TkError) as e:
        be.eigh_lowest(a, 4)  # EighKTooLarge
    assert e.value.status == tk_hip.TK_ERR_INVALID_ARG


@pytest.mark.parametrize("m,n", [(3, 2), (1, 1), (5, 5), (200, 150), (70, 130), (40, 33), (300, 64)])
def test_qr(be, m, n):
    print(f"case qr {m}x{n}", flush=True)
    if (m, n) == (3, 2):
        a = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])  # faer.rs:866-893
    else:
        a = np.random.default_rng(m * 7 + n).standard_normal((m, n))
    q, r = be.qr(a)
    k = min(m, n)
    assert q.shape == (m, k) and r.shape == (k, n)
    assert np.abs(q @ r - a).max() < 1e-10 * max(1.0, np.abs(a).max())
    assert np.abs(q.T @ q - np.eye(k)).max() < 1e-12
    assert np.abs(np.tril(r, -1)).max() == 0.0
    _, rl = np.linalg.qr(a)
    assert np.abs(np.abs(np.diag(r)) - np.abs(np.diag(rl))).max() < 1e-10 * np.abs(a).max()


def test_qr_rank_deficient(be):
    print("case qr rank deficient", flush=True)
    rng = np.random.default_rng(2)
    b = rng.standard_normal((50, 10))
    a = np.concatenate([b, b[:, :5], np.zeros((50, 3))], axis=1)
    q, r = be.qr(a)
    assert np.abs(q @ r - a).max() < 1e-10
    assert np.abs(q.T @ q - np.eye(18)).max() < 1e-12


def test_regularized_inverse(be):
    print("case regularized inverse", flush=True)
    # faer.rs:895-930
    eye = np.eye(3)
    r = be.regularized_svd_inverse(np.array([5.0, 3.0, 1.0]), eye, eye, 1e-10)
    assert abs(r[0, 0] - 0.2) < 1e-8 and abs(r[1, 1] - 1 / 3) < 1e-8 and abs(r[2, 2] - 1.0) < 1e-8
    r = be.regularized_svd_inverse(np.array([0.0]), np.eye(1), np.eye(1), 1e-8)
    assert abs(r[0, 0]) < 1e-12 and np.isfinite(r[0, 0])
    rng = np.random.default_rng(4)
    u, vt = _orth(rng, 6, 4), _orth(rng, 5, 4).T
    s = np.array([2.0, 1.0, 0.1, 1e-9])
    d = 1e-6
    want = vt.T @ np.diag(s / (s * s + d * d)) @ u.T
    assert np.abs(be.regularized_svd_inverse(s, u, vt, d) - want).max() < 1e-9 * np.abs(want).max()


def _chain(ctx, n=24):
    mpo = tk_hip.chain_mpo("heisenberg", n)
    bonds, data = tk_hip.initial_mps(tk_hip.SYM_U1, [m.phys for m in mpo], (0, 0), 16, 0)
    d = tk_hip.Dmrg(ctx, mpo, tk_hip.SYM_U1, schedule=(48,), davidson_tol=1e-10, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    return d


def test_cancel_flag(ctx):
    """run_with_cancel_flag (sweep/mod.rs:274-291): a raised flag stops the
    sweep before its next step with TK_ERR_CANCELLED (DmrgError::Cancelled);
    the engine resumes after prepare()."""
    print("case cancel flag", flush=True)
    d = _chain(ctx)
    d.cancel()
    with pytest.raises(tk_hip.TkError) as e:
        d.sweep()
    assert e.value.status == tk_hip.TK_ERR_CANCELLED
    with pytest.raises(tk_hip.TkError):
        d.sweep()  # not prepared again
    d.cancel(False)
    d.prepare()
    t0 = time.perf_counter()
    e_full = d.sweep()["energy"]
    t_sweep = time.perf_counter() - t0
    # from another thread, part-way through a sweep
    d.prepare()
    timer = threading.Timer(0.3 * t_sweep, d.cancel)
    timer.start()
    try:
        with pytest.raises(tk_hip.TkError) as e:
            for _ in range(50):
                d.sweep()
        assert e.value.status == tk_hip.TK_ERR_CANCELLED
    finally:
        timer.cancel()
    d.cancel(False)
    d.prepare()
    for _ in range(3):
        s = d.sweep()
    assert s["energy"] <= e_full + 1e-9
    d.close()


def _ulps(q, ref):
    """|q - ref| in units of the last place of ref (ref finite)."""
    return np.abs(q - ref) / np.spacing(np.abs(ref))


def test_fast_quotient_ulps(ctx):
    """The quotient of the eigensolver's Sturm counts and inverse iteration
    (hardware reciprocal of the divisor, one Newton step, product with the
    numerator) against IEEE division (numpy, correctly rounded): operands
    over the normal range, divisors down to the Sturm recurrence's floor
    (pivmin >= DBL_MIN), and numerators whose quotient is subnormal."""
    print("case fast quotient ulps", flush=True)
    rng = np.random.default_rng(7)
    tiny = np.finfo(np.float64).tiny
    sgn = lambda k: rng.choice([-1.0, 1.0], k)
    n = 200_000
    a = rng.uniform(1, 2, n) * np.exp2(rng.integers(-500, 500, n)) * sgn(n)
    b = rng.uniform(1, 2, n) * np.exp2(rng.integers(-500, 500, n)) * sgn(n)
    m = 50_000
    # divisors within a factor 1e3 of DBL_MIN, numerators keeping the quotient normal
    b2 = rng.uniform(1, 1e3, m) * tiny * sgn(m)
    a2 = rng.uniform(1, 2, m) * np.exp2(rng.integers(-1000, -30, m)) * sgn(m)
    # quotients below DBL_MIN (subnormal results)
    b3 = rng.uniform(1, 2, m) * sgn(m)
    a3 = rng.uniform(0, 1, m) * tiny * sgn(m)
    cases = {"normal": (a, b), "divisor near DBL_MIN": (a2, b2), "subnormal quotient": (a3, b3)}
    worst = {}
    for name, (x, y) in cases.items():
        dx, dy, dq = ctx.to_device(x), ctx.to_device(y), ctx.empty(len(x))
        tk_hip._check(ctx.L.tk_fast_quotient(ctx.h, tk_hip.C.c_void_p(dx.ptr), tk_hip.C.c_void_p(dy.ptr),
                                             tk_hip.C.c_void_p(dq.ptr), len(x)), ctx.h)
        q = dq.get()
        for d in (dx, dy, dq):
            d.free()
        ref = x / y
        assert np.all(np.isfinite(q)), name
        worst[name] = float(_ulps(q, ref).max())
    print("fast quotient, max ulps from IEEE division:", worst, flush=True)
    # one Newton step on the hardware reciprocal: 18 / 15 / 5 ulps measured;
    # the bounds leave a margin and fail by orders of magnitude without the step
    assert worst["normal"] <= 32.0 and worst["divisor near DBL_MIN"] <= 32.0
    assert worst["subnormal quotient"] <= 8.0
