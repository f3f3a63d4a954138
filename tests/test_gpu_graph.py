"""H_eff*psi replayed from a HIP graph against the eager apply.

The graph holds the same kernels with the same work lists, so every replay
equals the eager product bit for bit.  Replays re-point x and y between
calls and alternate with eager launches and host synchronisations on the same
stream; each is checked against its own eager result.
"""
import numpy as np
import pytest

import tk_hip
from heff_cases import TINY_UU_L, TINY_UU_R, random_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["tiny", "c2"])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_graph_apply_equals_eager(ctx, case, prec):
    print(f"case graph {case} {prec}", flush=True)
    p = random_problem("hubbard", 5, TINY_UU_L, TINY_UU_R) if case == "tiny" else tk_hip.c2_problem(seed=2024)
    W = p["W"]
    L, R = ctx.to_device(p["L"]), ctx.to_device(p["R"])
    plan = tk_hip.HeffPlan(ctx, prec, p["sym"], p["bond_l"], W.wl, L, p["bond_r"], W.wr, R, W, W)
    n = len(p["x"])
    rng = np.random.default_rng(1)
    xs = [ctx.to_device(rng.standard_normal(n)) for _ in range(3)]
    ys = [ctx.empty(n) for _ in range(4)]
    want = []
    for x in xs:
        plan.apply(x, ys[3])
        ctx.sync()
        want.append(ys[3].get())
    # replays: new x and y pointers each time, eager work and syncs between
    order = [(0, 0), (1, 1), (1, 2), (2, 0), (0, 1), (2, 2)]
    for rep, (i, j) in enumerate(order):
        plan.apply_graph(xs[i], ys[j])
        if rep % 2:
            plan.apply(xs[(i + 1) % 3], ys[3])
        ctx.sync()
        got = ys[j].get()
        assert np.array_equal(got, want[i]), (rep, i, j, np.abs(got - want[i]).max())
    plan.close()
    for a in xs + ys + [L, R]:
        a.free()
