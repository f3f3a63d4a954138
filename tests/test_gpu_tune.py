"""The contraction tiling autotuner on the GPU.

A context with tuning on measures every candidate tiling of an H_eff plan's
two stages at the plan's first apply and caches the fastest per stage
signature; a second plan with the same signatures reuses the cache; every
tiling gives the FP64 oracle's product (1e-12) and the default tiling's to
rounding; the cache survives a save / clear / load round trip.
"""
import os

import numpy as np
import pytest

import oracle
import tk_hip

pytestmark = pytest.mark.gpu


def _plan(ctx, p, prec):
    L, R = ctx.to_device(p["L"]), ctx.to_device(p["R"])
    W = p["W"]
    return tk_hip.HeffPlan(ctx, prec, p["sym"], p["bond_l"], W.wl, L, p["bond_r"], W.wr, R, W, W), (L, R)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_tuned_plan_matches_default_and_oracle(ctx, tmp_path, prec):
    print(f"case tune {prec}", flush=True)
    p = tk_hip.c2_problem(seed=2024)
    ref = oracle.heff_apply(p, p["x"])
    tol = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-4}[prec]
    # default tiling on the shared context
    pl0, keep0 = _plan(ctx, p, prec)
    x0, y0 = ctx.to_device(p["x"]), ctx.empty(len(p["x"]))
    pl0.apply(x0, y0)
    y_def = y0.get()
    assert pl0.tiling(1) == (1, 48) and pl0.tiling(3) == (1, 48)
    # a tuning context
    tc = tk_hip.Context(0)
    tc.tune(True)
    n_before = len(tk_hip.tune_table())
    pl, keep = _plan(tc, p, prec)
    x, y = tc.to_device(p["x"]), tc.empty(len(p["x"]))
    pl.apply(x, y)
    got = y.get()
    table = tk_hip.tune_table()
    assert len(table) >= n_before  # this precision's stages are measured now (or were already)
    kinds = {e["wg_kind"] for e in table}
    assert kinds <= {0, 1, 2}
    for stage in (1, 3):
        k, t = pl.tiling(stage)
        assert k in (0, 1, 2) and t in (32, 48, 64)
        if prec == "bf16x3":
            assert k != 2
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    assert np.abs(got - y_def).max() <= (1e-13 if prec == "fp64" else tol) * np.abs(ref).max()
    # the same signatures again: cached, the same tiling, no new entries
    pl2, keep2 = _plan(tc, p, prec)
    y2 = tc.empty(len(p["x"]))
    pl2.apply(x, y2)
    assert len(tk_hip.tune_table()) == len(table)
    assert pl2.tiling(1) == pl.tiling(1) and pl2.tiling(3) == pl.tiling(3)
    assert np.array_equal(y2.get(), got)
    # save / clear / load
    path = os.path.join(str(tmp_path), "tune.txt")
    tk_hip.tune_save(path)
    tk_hip.tune_clear()
    assert tk_hip.tune_table() == []
    tk_hip.tune_load(path)
    assert sorted((e["signature"], e["wg_kind"], e["tile_min"]) for e in tk_hip.tune_table()) == \
        sorted((e["signature"], e["wg_kind"], e["tile_min"]) for e in table)
    for q in (pl, pl2, pl0):
        q.close()
    tc.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_every_tiling_matches_oracle(tmp_path, prec):
    """Each workgroup form, forced through a loaded cache entry for the
    plan's stage signatures, gives the oracle's product: wavefront tiles,
    32 x 128 and 64 x 64 (at two thresholds).  Split-BF16 has only the
    wavefront and 32 x 128 forms and runs the 64 x 64 form's outputs as
    wavefront tiles."""
    print(f"case forced tilings {prec}", flush=True)
    p = tk_hip.c2_problem(seed=2025)
    ref = oracle.heff_apply(p, p["x"])
    tol = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-4}[prec]
    tc = tk_hip.Context(0)
    tc.tune(True)
    pl, keep = _plan(tc, p, prec)
    x, y = tc.to_device(p["x"]), tc.empty(len(p["x"]))
    pl.apply(x, y)
    sigs = sorted({e["signature"] for e in tk_hip.tune_table()})
    pl.close()
    path = os.path.join(str(tmp_path), "forced.txt")
    try:
        for kind, tmin in ((0, 48), (1, 48), (2, 32), (2, 64)):
            with open(path, "w") as fh:
                for s in sigs:
                    fh.write(f"0 {s} {kind} {tmin} 1 1\n")
            tk_hip.tune_clear()
            tk_hip.tune_load(path)
            q, keep_q = _plan(tc, p, prec)
            assert q.tiling(1) == (kind, tmin) and q.tiling(3) == (kind, tmin)
            yq = tc.empty(len(p["x"]))
            q.apply(x, yq)
            got = yq.get()
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"  kind {kind} tile_min {tmin}: rel err {err:.2e}", flush=True)
            assert err <= tol
            q.close()
    finally:
        tk_hip.tune_clear()
        tc.close()


def test_tuned_sweep_matches_untuned(ctx):
    """A DMRG run with tuning on reaches the untuned run's energies."""
    print("case tune dmrg", flush=True)
    n = 12
    mpo = tk_hip.chain_mpo("hubbard", n, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (6, 6), 16, 0)
    out = []
    for tune in (False, True):
        c = tk_hip.Context(0)
        c.tune(tune)
        d = tk_hip.Dmrg(c, mpo, sym, schedule=(24, 48), davidson_tol=1e-12, svd_cutoff=1e-10)
        d.set_state(bonds, data)
        d.prepare()
        out.append([d.sweep()["energy"] for _ in range(3)])
        d.close()
        c.close()
    assert abs(out[0][-1] - out[1][-1]) <= 1e-10
