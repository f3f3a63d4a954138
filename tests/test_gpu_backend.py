"""LinAlgBackend / SparseLinAlgBackend trait methods on the GPU through the C ABI,
on the reference's own known answers and LCG goldens:

* gemm identity and alpha/beta (tk-linalg/src/device/faer.rs:752-793);
* block_gemm vs dense GEMM (tk-linalg/tests/block_gemm_realistic.rs:91-196) and
  its output table vs the oracle's (bit-exact);
* spmv vs dense GEMV (tk-linalg/tests/spmv_test.rs, seeds 42/137, 77/200, 99/300);
* svd_truncated reconstruction / max_rank (faer.rs:795-844) and vs LAPACK;
* eigh 2 - sqrt(2) (faer.rs:846-864); Davidson known answers (eigensolver/mod.rs:709-801).
"""
import math

import numpy as np
import pytest

import oracle
import tk_hip
from conftest import lcg_next

pytestmark = pytest.mark.gpu
U1 = tk_hip.SYM_U1


@pytest.fixture(scope="module")
def be(ctx):
    return tk_hip.HipBackend(ctx)


def test_gemm_identity(be):
    c = be.gemm(1.0, np.array([[1.0, 2], [3, 4]]), np.eye(2), 0.0, np.zeros((2, 2)))
    assert np.allclose(c.ravel(), [1, 2, 3, 4], atol=1e-12)


def test_gemm_alpha_beta(be):
    c = be.gemm(2.0, np.array([[1.0, 2], [3, 4]]), np.eye(2), 3.0, np.array([[10.0, 20], [30, 40]]))
    assert np.allclose(c.ravel(), [32, 64, 96, 128], atol=1e-12)


def test_gemm_strided_views(be, ctx):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((37, 23))
    b = rng.standard_normal((23, 41))
    # A passed as a column-major view of a^T storage, B row-major, C column-major
    at = np.ascontiguousarray(a.T)
    da, db = ctx.to_device(at), ctx.to_device(b)
    dc = ctx.empty(37 * 41)
    be.gemm_strided(37, 41, 23, 1.0, da, 1, 37, db, 41, 1, 0.0, dc, 1, 37)
    c = dc.get().reshape(41, 37).T
    assert np.abs(c - a @ b).max() < 1e-12


def _lcg_matrix(rows, cols, flux, seed):
    rs, cs = sorted(rows), sorted(cols)
    state = seed
    blocks = {}
    for rq, rd in rs:
        for cq, cd in cs:
            if rq - cq != flux:
                continue
            vals = []
            for _ in range(rd * cd):
                state, v = lcg_next(state)
                vals.append(v)
            blocks[oracle.pack_key(U1, [rq, cq])] = np.array(vals).reshape(rd, cd)
    flat = np.concatenate([blocks[k].ravel() for k in sorted(blocks)])
    ro = dict(zip([q for q, _ in rs], np.cumsum([0] + [d for _, d in rs])[:-1]))
    co = dict(zip([q for q, _ in cs], np.cumsum([0] + [d for _, d in cs])[:-1]))
    dense = np.zeros((sum(d for _, d in rs), sum(d for _, d in cs)))
    for k, b in blocks.items():
        (rq, _), (cq, _) = oracle.unpack_key(U1, k, 2)
        dense[ro[rq]:ro[rq] + b.shape[0], co[cq]:co[cq] + b.shape[1]] = b
    return flat, dense


def _dense_from(keys, offs, data, rows, cols):
    rs, cs = sorted(rows), sorted(cols)
    ro = dict(zip([q for q, _ in rs], np.cumsum([0] + [d for _, d in rs])[:-1]))
    co = dict(zip([q for q, _ in cs], np.cumsum([0] + [d for _, d in cs])[:-1]))
    rd, cd = dict(rs), dict(cs)
    out = np.zeros((sum(rd.values()), sum(cd.values())))
    for k, o in zip(keys, offs):
        (rq, _), (cq, _) = oracle.unpack_key(U1, k, 2)
        out[ro[rq]:ro[rq] + rd[rq], co[cq]:co[cq] + cd[cq]] = data[o:o + rd[rq] * cd[cq]].reshape(rd[rq], cd[cq])
    return out


@pytest.mark.parametrize("rows,cols,fa,fb,sa,sb", [
    ([(-1, 4), (0, 8), (1, 4)], [(-1, 3), (0, 6), (1, 3)], 0, 0, 42, 137),
    ([(0, 3), (1, 5)], [(-1, 2), (0, 4)], 1, -1, 99, 200),
])
def test_block_gemm_lcg_golden(be, rows, cols, fa, fb, sa, sb):
    a_flat, a_dense = _lcg_matrix(rows, cols, fa, sa)
    b_flat, b_dense = _lcg_matrix(cols, rows, fb, sb)
    keys, offs, c = be.block_gemm(U1, rows, cols, fa, a_flat, cols, rows, fb, b_flat)
    ok, oo, oc = oracle.block_gemm(U1, rows, cols, (fa, 0), a_flat, cols, rows, (fb, 0), b_flat)
    assert np.array_equal(keys, ok) and np.array_equal(offs, oo)  # output table bit-exact
    assert np.abs(c - oc).max() < 1e-13
    assert np.abs(_dense_from(keys, offs, c, rows, rows) - a_dense @ b_dense).max() < 1e-10


@pytest.mark.parametrize("rows,cols,flux,sa,sx", [
    ([(-1, 2), (0, 3), (1, 2)], [(-1, 2), (0, 3), (1, 2)], 0, 42, 137),
    ([(-2, 3), (0, 5), (2, 3)], [(-2, 2), (0, 4), (2, 4)], 0, 77, 200),
    ([(0, 3), (1, 4)], [(-1, 2), (0, 5)], 1, 99, 300),
])
def test_spmv_lcg_golden(be, rows, cols, flux, sa, sx):
    a_flat, a_dense = _lcg_matrix(rows, cols, flux, sa)
    state = sx
    x = []
    for _ in range(sum(d for _, d in cols)):
        state, v = lcg_next(state)
        x.append(v)
    x = np.array(x)
    y = be.spmv(U1, rows, cols, flux, a_flat, x)
    assert np.abs(y - a_dense @ x).max() < 1e-12


def test_svd_reconstruction(be):
    a = np.array([[1.0, 2], [3, 4], [5, 6]])
    u, s, vt, te = be.svd_truncated(a, 2, 0.0)
    assert len(s) == 2 and s[0] >= s[1]
    assert np.abs(u @ np.diag(s) @ vt - a).max() < 1e-10
    u, s, vt, te = be.svd_truncated(a, 1, 0.0)
    assert len(s) == 1 and te > 0.0
    sv = np.linalg.svd(a, compute_uv=False)
    assert abs(te - sv[1] ** 2) < 1e-10


@pytest.mark.parametrize("shape", [(40, 25), (25, 40), (64, 64)])
def test_svd_matches_lapack(be, shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    a = rng.standard_normal(shape)
    u, s, vt, te = be.svd_truncated(a, min(shape), 0.0)
    ref = np.linalg.svd(a, compute_uv=False)
    assert np.abs(s - ref).max() < 1e-10 * ref[0]
    assert np.abs(u @ np.diag(s) @ vt - a).max() < 1e-10
    assert np.abs(u.T @ u - np.eye(len(s))).max() < 1e-10


def test_eigh_lowest(be):
    w, v = be.eigh(np.array([[2.0, -1, 0], [-1, 2, -1], [0, -1, 2]]))
    assert abs(w[0] - (2 - math.sqrt(2))) < 1e-10 and w[0] <= w[1] <= w[2]


@pytest.mark.parametrize("n", [1, 2, 3, 50, 128, 129, 192, 193, 300, 700, 1100, 2048])
def test_eigh_matches_lapack(be, n):
    """Tridiagonalisation (one workgroup up to 192, column steps above),
    bisection, inverse iteration, CGS2 and back-transformation against
    LAPACK's eigenvalues; eigenvectors by residual and orthonormality."""
    rng = np.random.default_rng(1000 + n)
    x = rng.standard_normal((n, n))
    a = (x + x.T) / 2
    w, v = be.eigh(a)
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(ref).max()
    assert np.abs(w - ref).max() < 1e-12 * scale * max(1.0, math.sqrt(n) / 4)
    assert np.abs(a @ v - v * w).max() < 1e-11 * scale
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12 * max(1.0, n / 50)


@pytest.mark.parametrize("n", [40, 300])
def test_eigh_degenerate_spectrum(be, n):
    """Exactly repeated eigenvalues (multiplets of 1, 2, 3 and 5): the
    eigenvectors must stay orthonormal and span each eigenspace."""
    rng = np.random.default_rng(7 * n)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    mult = [1, 2, 3, 5]
    lam = []
    level = 0
    while len(lam) < n:
        lam += [float(level)] * mult[level % 4]
        level += 1
    lam = np.array(lam[:n])
    a = (q * lam) @ q.T
    a = (a + a.T) / 2
    w, v = be.eigh(a)
    assert np.abs(w - np.sort(lam)).max() < 1e-11 * lam.max()
    assert np.abs(a @ v - v * w).max() < 1e-10 * lam.max()
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-11


class _Dev:
    def __init__(self, p):
        self.ptr = p


def _dense_matvec(be, ctx, h):
    n = h.shape[0]
    dh = ctx.to_device(h)

    def mv(x, y):
        be.gemm_strided(n, 1, n, 1.0, dh, n, 1, _Dev(x), 1, 1, 0.0, _Dev(y), 1, 1)
    return mv, dh


def test_davidson_known_answers(be, ctx):
    d = np.array([3.0, 1, 4, 1, 5])
    x0 = np.array([((i * 7 + 13) % 97) / 97.0 - 0.5 for i in range(5)])
    mv, keep = _dense_matvec(be, ctx, np.diag(d))
    e, x, rn, nmv = be.davidson(mv, 5, x0, diag=d)
    assert abs(e - 1.0) < 1e-6
    mv, keep = _dense_matvec(be, ctx, np.array([[2.0, -1], [-1, 2]]))
    e, x, rn, nmv = be.davidson(mv, 2, x0[:2], diag=np.array([2.0, 2.0]), tol=1e-8)
    assert abs(e - 1.0) < 1e-6
    mv, keep = _dense_matvec(be, ctx, np.diag([5.0, 2, 8, 1, 3]))
    e, x, rn, nmv = be.davidson(mv, 5, x0, tol=1e-8)
    assert abs(e - 1.0) < 1e-6


def test_davidson_random_vs_lapack(be, ctx):
    rng = np.random.default_rng(11)
    a = rng.standard_normal((48, 48))
    h = a + a.T
    mv, keep = _dense_matvec(be, ctx, h)
    e, x, rn, nmv = be.davidson(mv, 48, rng.standard_normal(48), diag=np.diag(h).copy(), tol=1e-11)
    assert abs(e - np.linalg.eigvalsh(h)[0]) < 1e-10
    assert np.linalg.norm(h @ x - e * x) < 1e-9


def test_lanczos_known_answers(be, ctx):
    """LanczosSolver tests (eigensolver/mod.rs:660-700): diag(3,1,4,1,5) and
    [[2,-1],[-1,2]] (max_krylov 10) from the reference start vector; the
    eigenvector is checked by its residual (the reference leaves it unset)."""
    x0 = np.array([((i * 7 + 13) % 97) / 97.0 - 0.5 for i in range(5)])
    h = np.diag([3.0, 1, 4, 1, 5])
    mv, keep = _dense_matvec(be, ctx, h)
    e, x, rn, nmv = be.lanczos(mv, 5, x0)
    assert abs(e - 1.0) < 1e-10
    assert np.linalg.norm(h @ x - e * x) < 1e-9 and abs(np.linalg.norm(x) - 1) < 1e-12
    h = np.array([[2.0, -1], [-1, 2]])
    mv, keep = _dense_matvec(be, ctx, h)
    e, x, rn, nmv = be.lanczos(mv, 2, x0[:2], max_krylov=10)
    assert abs(e - 1.0) < 1e-10
    assert np.linalg.norm(h @ x - e * x) < 1e-9


def _lcg_diag(dim, seed):
    """Random diagonal of proptest_dmrg.rs:18-30 (LCG, state = seed + 42)."""
    state = (seed + 42) % 2 ** 64
    out = []
    for _ in range(dim):
        state = (state * 6364136223846793005 + 1) % 2 ** 64
        v = (state >> 33) & 0xFFFFFFFF
        v = v - 2 ** 32 if v >= 2 ** 31 else v
        out.append(v / float(1 << 30))
    return np.array(out)


@pytest.mark.parametrize("dim,seed", [(3, 0), (5, 17), (8, 123), (13, 999), (20, 4242), (20, 9999)])
def test_lanczos_random_diagonal(be, ctx, dim, seed):
    """proptest lanczos_eigenvalue_accuracy (proptest_dmrg.rs:15-49): the
    minimum of an LCG diagonal, from the reference start vector."""
    d = _lcg_diag(dim, seed)
    x0 = np.array([((i * 7 + 13) % 97) / 97.0 - 0.5 for i in range(dim)])
    mv, keep = _dense_matvec(be, ctx, np.diag(d))
    e, x, rn, nmv = be.lanczos(mv, dim, x0)
    assert np.isfinite(e) and abs(e - d.min()) < 1e-9


def test_lanczos_random_vs_lapack(be, ctx):
    rng = np.random.default_rng(12)
    a = rng.standard_normal((200, 200))
    h = a + a.T
    mv, keep = _dense_matvec(be, ctx, h)
    e, x, rn, nmv = be.lanczos(mv, 200, rng.standard_normal(200), tol=1e-11, max_krylov=60)
    assert abs(e - np.linalg.eigvalsh(h)[0]) < 1e-10
    assert np.linalg.norm(h @ x - e * x) < 1e-9 and rn < 1e-9
