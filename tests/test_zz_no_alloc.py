"""tk_gemm_noalloc (no device allocation; work list in a static device array)
against tk_gemm through the C ABI: bit-identical results on two calls in a
row, the second smaller, over plain, transposed-view and beta-scaled
operands; outputs beyond the array's size are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

import tk_hip

pytestmark = pytest.mark.gpu


def _call(ctx, fn, m, n, k, alpha, a, rsa, csa, b, rsb, csb, beta, c, rsc, csc):
    st = getattr(ctx.L, fn)(ctx.h, m, n, k, alpha, C.c_void_p(a.ptr if a else None), rsa, csa,
                            C.c_void_p(b.ptr if b else None), rsb, csb, beta, C.c_void_p(c.ptr if c else None),
                            rsc, csc)
    return st


def _both(ctx, m, n, k, alpha, beta, ta, tb, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((k, m) if ta else (m, k))
    b = rng.standard_normal((n, k) if tb else (k, n))
    c0 = rng.standard_normal((m, n))
    da, db = ctx.to_device(a), ctx.to_device(b)
    rsa, csa = (1, m) if ta else (k, 1)
    rsb, csb = (1, k) if tb else (n, 1)
    out = {}
    for fn in ("tk_gemm", "tk_gemm_noalloc"):
        dc = ctx.to_device(c0)
        assert _call(ctx, fn, m, n, k, alpha, da, rsa, csa, db, rsb, csb, beta, dc, n, 1) == tk_hip.TK_OK, \
            tk_hip.lib().tk_last_error(ctx.h)
        out[fn] = dc.get().reshape(m, n)
        dc.free()
    da.free()
    db.free()
    want = alpha * (a.T if ta else a) @ (b.T if tb else b) + beta * c0
    return out["tk_gemm"], out["tk_gemm_noalloc"], want


@pytest.mark.parametrize("ta,tb,beta", [(False, False, 0.0), (True, False, 0.5), (False, True, -1.0)])
def test_gemm_noalloc_equals_gemm(ctx, ta, tb, beta):
    print(f"case gemm_noalloc ta={ta} tb={tb} beta={beta}", flush=True)
    for (m, n, k), seed in (((300, 257, 190), 1), ((37, 70, 5), 2)):  # the second call smaller
        ref, got, want = _both(ctx, m, n, k, 1.5, beta, ta, tb, seed)
        assert np.array_equal(ref, got)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_gemm_noalloc_beta_only_and_empty(ctx):
    print("case gemm_noalloc k=0 and empty", flush=True)
    ref, got, want = _both(ctx, 50, 40, 0, 1.0, 2.0, False, False, 3)
    assert np.array_equal(ref, got) and np.abs(got - want).max() == 0.0
    assert _call(ctx, "tk_gemm_noalloc", 0, 5, 3, 1.0, None, 3, 1, None, 5, 1, 0.0, None, 5, 1) == tk_hip.TK_OK


def test_gemm_noalloc_rejects_oversize(ctx):
    print("case gemm_noalloc oversize", flush=True)
    # 625 x 625 tiles = 6.25 MB of work list > the 4 MiB array; rejected on the host before any launch
    st = _call(ctx, "tk_gemm_noalloc", 20000, 20000, 1, 1.0, None, 1, 1, None, 20000, 1, 0.0, None, 20000, 1)
    assert st == tk_hip.TK_ERR_INVALID_ARG
    assert "static scratch" in tk_hip.lib().tk_last_error(ctx.h).decode()
