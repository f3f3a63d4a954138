"""The XCD-local cooperative column steps of the tridiagonalisation against
the one-launch-per-column steps and LAPACK, through the C ABI.

Both paths run in one process (TK_TRIDIAG_COOP=0 selects the per-column
launches for a call); sizes cover both register variants (n <= 512, n <= 1024),
the hand-over to the single-workgroup tail (192 columns), matrices above 1024
rows (left to the per-column launches inside the same batch) and batches that
need more than one cooperative launch (more rows than the XCDs' registers hold).
"""
import os

import numpy as np
import pytest

import tk_hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be(ctx):
    return tk_hip.HipBackend(ctx)


def _coop_off(fn):
    os.environ["TK_TRIDIAG_COOP"] = "0"
    try:
        return fn()
    finally:
        del os.environ["TK_TRIDIAG_COOP"]


def _sigma_ok(s, s_true, smax):
    """Two Gram levels: sigma >= 3.2e-5 sigma_max from an eigenvalue of X X^T
    accurate to ~eps sigma_max^2; below, ~eps sigma_max in sigma itself."""
    tol = np.where(s_true >= 3.2e-5 * smax, 1e-14 * smax ** 2 / s_true, 1e-13 * smax)
    return bool(np.all(np.abs(s - s_true) <= tol))


def _sym(rng, n):
    a = rng.standard_normal((n, n))
    return (a + a.T) / 2


@pytest.mark.parametrize("n", [193, 330, 512, 513, 700, 1024, 1030])
def test_eigh_coop_matches_per_column_and_lapack(be, ctx, n):
    print(f"case coop eigh n={n}", flush=True)
    rng = np.random.default_rng(1000 + n)
    a = _sym(rng, n)
    ctx.timing(True)
    w, v = be.eigh(a)
    t = ctx.timing_read()
    ctx.timing(False)
    # the cooperative kernel ran for every size it holds
    assert ("tridiag_coop_fp64" in t and t["tridiag_coop_fp64"]["bytes"] > 0) == (n <= 1024)
    w0, v0 = _coop_off(lambda: be.eigh(a))
    ref = np.linalg.eigvalsh(a)
    amax = np.abs(a).max()
    assert np.abs(w - ref).max() < 1e-13 * n * amax
    assert np.abs(w - w0).max() < 1e-13 * n * amax
    assert np.abs(a @ v - v * w).max() < 1e-12 * amax * np.sqrt(n)
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12


def test_block_svd_coop_many_large_blocks(be, ctx):
    """Twelve blocks of 600-1000 rows and a few small ones: more rows than one
    cooperative launch places, so the batch runs in several; kept values and
    subspaces equal the per-column path's and LAPACK's."""
    print("case coop block svd many blocks", flush=True)
    rng = np.random.default_rng(77)
    blocks = []
    for q in range(12):
        m = int(rng.integers(600, 1001))
        n = int(rng.integers(500, 1001))
        k = min(m, n)
        s = np.sort(10.0 ** rng.uniform(-6, 0, k))[::-1]
        qa, _ = np.linalg.qr(rng.standard_normal((m, k)))
        qb, _ = np.linalg.qr(rng.standard_normal((n, k)))
        blocks.append((qa * s) @ qb.T)
    for m, n in [(40, 30), (150, 200), (250, 90)]:
        blocks.append(rng.standard_normal((m, n)))
    out, nk, nd, err = be.block_svd_truncate(blocks, 10 ** 6, 1e-12)
    out0, nk0, nd0, err0 = _coop_off(lambda: be.block_svd_truncate(blocks, 10 ** 6, 1e-12))
    assert nk == nk0 and nd == nd0
    sig = [np.linalg.svd(b, compute_uv=False) for b in blocks]
    smax = max(x[0] for x in sig)
    for b, (u, s, c), (u0, s0, c0), sl in zip(blocks, out, out0, sig):
        assert len(s) == len(s0) == len(sl)
        assert _sigma_ok(s, sl, smax) and _sigma_ok(s0, sl, smax)
        assert np.abs(u.T @ u - np.eye(len(s))).max() < 1e-10
        assert np.abs(u @ c - b).max() <= 1e-10 * smax
