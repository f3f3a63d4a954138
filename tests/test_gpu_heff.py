"""H_eff*psi on the GPU through the C ABI vs the FP64 oracle.

Bars: sector tables bit-exact; FP64 within 1e-12 relative (max-norm over the
output); FP32 and split-BF16 within their stated tolerances; results are
deterministic (bit-identical on repeat).
"""
import numpy as np
import pytest

import oracle
import tk_hip
from heff_cases import TINY_U1_L, TINY_U1_R, TINY_UU_L, TINY_UU_R, random_problem

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-4}


def _plan(ctx, p, prec, Wi=None, Wj=None):
    Wi = Wi or p["W"]
    Wj = Wj or p["W"]
    L = ctx.to_device(p["L"])
    R = ctx.to_device(p["R"])
    plan = tk_hip.HeffPlan(ctx, prec, p["sym"], p["bond_l"], Wi.wl, L, p["bond_r"], Wj.wr, R, Wi, Wj)
    return plan, (L, R)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("kind,bl,br", [("hubbard", TINY_UU_L, TINY_UU_R), ("heisenberg", TINY_U1_L, TINY_U1_R)])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_tiny_heff_matches_oracle(ctx, kind, bl, br, prec):
    print(f"case tiny {kind} {prec}", flush=True)
    p = random_problem(kind, 11, bl, br)
    plan, keep = _plan(ctx, p, prec)
    x = ctx.to_device(p["x"])
    y = ctx.empty(len(p["x"]))
    plan.apply(x, y)
    got = y.get()
    ref = oracle.heff_apply(p, p["x"])
    assert _rel(got, ref) <= TOL[prec], (prec, _rel(got, ref))


@pytest.fixture(scope="module")
def c2():
    return tk_hip.c2_problem(seed=2024)


def test_c2_psi_table_bit_exact(ctx, c2):
    print("case c2 table", flush=True)
    plan, keep = _plan(ctx, c2, "fp64")
    keys, offs, tot = plan.psi_table()
    W = c2["W"]
    ok, oo, ot = oracle.block_table(c2["sym"], [W.phys, W.phys, c2["bond_l"], c2["bond_r"]], [1, 1, 1, 0])
    assert np.array_equal(keys, ok) and np.array_equal(offs, oo) and tot == ot
    st = plan.stats()
    assert st["dim"] == tot and 1800 <= st["psi_blocks"] <= 2600


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_c2_heff_matches_oracle(ctx, c2, prec):
    print(f"case c2 heff {prec}", flush=True)
    plan, keep = _plan(ctx, c2, prec)
    x = ctx.to_device(c2["x"])
    y = ctx.empty(len(c2["x"]))
    plan.apply(x, y)
    got = y.get()
    ref = oracle.heff_apply(c2, c2["x"])
    rel = _rel(got, ref)
    assert rel <= TOL[prec], (prec, rel)
    # relative 2-norm error too
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL[prec]


def test_c2_fp64_refinement_path_and_determinism(ctx, c2):
    print("case c2 refine/determinism", flush=True)
    p64, k1 = _plan(ctx, c2, "fp64")
    p32, k2 = _plan(ctx, c2, "fp32")
    x = ctx.to_device(c2["x"])
    y1, y2, y3 = ctx.empty(len(c2["x"])), ctx.empty(len(c2["x"])), ctx.empty(len(c2["x"]))
    p64.apply(x, y1)
    p64.apply(x, y2)
    p32.apply_fp64(x, y3)
    a, b, c = y1.get(), y2.get(), y3.get()
    assert np.array_equal(a, b)  # deterministic accumulation order
    assert np.array_equal(a, c)  # the refinement matvec of a low-precision plan is the FP64 kernel


def test_c2_diag_matches_oracle(ctx, c2):
    print("case c2 diag", flush=True)
    plan, keep = _plan(ctx, c2, "fp64")
    d = ctx.empty(len(c2["x"]))
    plan.diag(d)
    ref = oracle.heff_diag(c2)
    assert _rel(d.get(), ref) <= 1e-13


@pytest.fixture(scope="module")
def c4():
    p = tk_hip.c4_problem(seed=4096)
    p["y_ref"] = oracle.heff_apply(p, p["x"], Wi=p["W"], Wj=p["Wj"])
    return p


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_c4_cylinder_heff_matches_oracle(ctx, c4, prec):
    """Width-6 cylinder centre bond: MPO bond 20, U(1) bonds of 4096 states in
    31 skewed sectors (largest 406), 4.66 M psi entries."""
    print(f"case c4 heff {prec}", flush=True)
    plan, keep = _plan(ctx, c4, prec, c4["W"], c4["Wj"])
    keys, offs, tot = plan.psi_table()
    ok, oo, ot = oracle.block_table(c4["sym"], [c4["W"].phys, c4["Wj"].phys, c4["bond_l"], c4["bond_r"]],
                                    [1, 1, 1, 0])
    assert np.array_equal(keys, ok) and np.array_equal(offs, oo) and tot == ot == len(c4["x"])
    x = ctx.to_device(c4["x"])
    y = ctx.empty(len(c4["x"]))
    plan.apply(x, y)
    got = y.get()
    ref = c4["y_ref"]
    assert _rel(got, ref) <= TOL[prec], (prec, _rel(got, ref))
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= TOL[prec]
    plan.close()
    for b in keep + (x, y):
        b.free()


def test_heisenberg_two_site_singlet_gpu(ctx):
    print("case singlet", flush=True)
    Wl, Wr = tk_hip.heisenberg_w(1.0, "left"), tk_hip.heisenberg_w(1.0, "right")
    triv = [(0, 1)]
    p = {"bond_l": triv, "bond_r": triv, "sym": tk_hip.SYM_U1, "L": np.ones(1), "R": np.ones(1)}
    plan, keep = _plan(ctx, p, "fp64", Wl, Wr)
    n = plan.stats()["dim"]
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        x = ctx.to_device(e)
        y = ctx.empty(n)
        plan.apply(x, y)
        H[:, i] = y.get()
    assert np.allclose(H, H.T, atol=1e-15)
    assert abs(np.linalg.eigvalsh(H).min() + 0.75) < 1e-14
