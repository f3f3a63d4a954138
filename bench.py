#!/usr/bin/env python3
"""Benchmark of the block-sparse two-site DMRG hot path on one MI355X.

Workloads
  sweep  (default) two-site DMRG on the Hubbard chain L=64, t=1, U=4, half
         filling, Sz=0, U(1)xU(1): bond dimension 256 -> 2048 over 8 sweeps,
         run in FP64 and in mixed precision (by default split-BF16 products
         with FP64 Krylov vectors, SVD and refinement; --mixed fp32 for FP32
         MFMA) from the same seeded MPS; a step is one full sweep at D=2048.  The ramp sweeps (D < 2048) are never
         timed; --warmup untimed sweeps at D=2048 follow them, then --steps
         timed ones; value is the median timed mixed sweep.
  heff   one H_eff*psi on the U(1)xU(1) Hubbard microbenchmark bond
         (d=4, MPO bond 6, D=1024 on both sides, 145 sectors); a step is one
         matvec with inputs resident in HBM.
  heff-c3  one H_eff*psi on the centre bonds of the L=64 Hubbard ground
         state ramped to D=2048 (the timed sweeps' shapes), random data.
  heff-d2048 the same construction with D=2048 on both sides (the bond
         dimension of the timed sweeps); every precision timed.
  sweep-c4 two-site DMRG on the J1 Heisenberg 6 x 12 cylinder (U(1), Sz=0),
         D 256 -> 4096 over 5 sweeps, in FP64, FP32 and split-BF16 from the
         same seeded MPS; a step is the sweep at D=4096.
  heff-c4  one H_eff*psi at the centre bond of the J1 Heisenberg 6 x 12
         cylinder (snake order, MPO bond 20, U(1) D=4096 in 31 sectors);
         FP64, FP32 and split-BF16 plans compared against the FP64 output.
  copy   calibration copy of known size (for the HBM counters).

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is refused: this build
runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tensorkraft-design_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
METRIC = "DMRG sweep time, Hubbard L=64 D=2048, mixed vs FP64 GPU at |ΔE|≤10^-10"


def cpu_baseline_heff(p, budget_s=10.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    wi, wj = p["W"], p.get("Wj", p["W"])
    oracle.heff_apply(p, p["x"], Wi=wi, Wj=wj)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.heff_apply(p, p["x"], Wi=wi, Wj=wj)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 500:
            break
    cores = os.cpu_count()
    return {"value": 1e6 * el / n, "unit": "us/matvec", "cores": cores, "kind": "port",
            "sample": f"{n} FP64 H_eff*psi matvecs of the same workload on {cores} OpenMP threads"}


HEFF_WORKLOADS = {
    "heff": ("heff_c2", "H_eff*psi time, Hubbard U(1)xU(1) d=4 w=6 D=1024 (single matvec)"),
    "heff-d2048": ("heff_hubbard_uu_D2048", "H_eff*psi time, Hubbard U(1)xU(1) d=4 w=6 D=2048 (single matvec)"),
    "heff-c3": ("heff_c3_hubbard_L64_D2048_centre",
                "H_eff*psi time, Hubbard L=64 D=2048 ground state's centre bond (single matvec)"),
    "heff-c4": ("heff_c4_cylinder_6x12_D4096",
                "H_eff*psi time, J1 Heisenberg 6x12 cylinder U(1) d=2 w=20 D=4096 (single matvec)"),
}


def heff_problem_for(tk_hip, workload, args=None, ctx=None):
    if workload == "heff-c3":
        # the bonds either side of the centre pair of the L=64 chain's ground
        # state ramped to D=2048 (the timed sweeps' largest H_eff); random
        # environments and psi of that block structure
        def log(msg):
            print(msg, file=sys.stderr, flush=True)

        d, mpo, sym, _, _ = _ground_state(tk_hip, ctx, args, log)
        n = args.sites
        bl, br = d.bond(n // 2 - 1), d.bond(n // 2 + 1)
        d.close()
        log("centre bonds " + json.dumps({"left": bl, "right": br}))
        p = tk_hip.heff_problem(bl, br, mpo[n // 2], seed=2048, sym=sym)
        desc = {"bond_dim": sum(x for _, x in bl), "sectors": len(bl), "mpo_bond_dim": 6,
                "bond": [n // 2 - 1, n // 2], "sector_dim_max": max(x for _, x in bl)}
    elif workload == "heff-c4":
        p = tk_hip.c4_problem(seed=4096)
        desc = {"bond_dim": 4096, "sectors": len(p["bond_l"]), "mpo_bond_dim": int(p["W"].w.shape[3]),
                "lattice": "cylinder 6 x 12, snake order", "bond": [p["site"], p["site"] + 1]}
    elif workload == "heff-d2048":
        p = tk_hip.c2_problem(seed=2048, total=2048)
        desc = {"bond_dim": 2048, "sectors": len(p["bond_l"]), "mpo_bond_dim": 6}
    else:
        p = tk_hip.c2_problem(seed=2024)
        desc = {"bond_dim": 1024, "sectors": len(p["bond_l"]), "mpo_bond_dim": 6}
    return p, desc


def bench_heff(args):
    """One H_eff*psi per step at the chosen precision; every precision's plan
    is also timed (HIP events) and compared with the FP64 plan's output."""
    import numpy as np

    import tk_hip

    ctx = tk_hip.Context(0)
    p, desc = heff_problem_for(tk_hip, args.workload, args, ctx)
    ctx.tune(not args.no_tune)
    L, R, x = ctx.to_device(p["L"]), ctx.to_device(p["R"]), ctx.to_device(p["x"])
    y = ctx.empty(len(p["x"]))
    Wi, Wj = p["W"], p.get("Wj", p["W"])

    def make(prec):
        return tk_hip.HeffPlan(ctx, prec, p["sym"], p["bond_l"], Wi.wl, L, p["bond_r"], Wj.wr, R, Wi, Wj)

    compare, y64 = {}, None
    for prec in ["fp64"] + [q for q in ("fp32", "bf16x3") if q != "fp64"]:
        pl = make(prec)
        pl.apply(x, y)
        ctx.sync()
        yy = y.get()
        if prec == "fp64":
            y64 = yy
        us_a, _, _ = pl.time(x, y, 20)
        compare[prec] = {"us": round(us_a, 3),
                         "tflops": round(pl.stats()["flops"] / (us_a * 1e-6) / 1e12, 3),
                         "rel_err_vs_fp64": float(np.linalg.norm(yy - y64) / np.linalg.norm(y64))}
        pl.close()
    plan = make(args.prec)
    st = plan.stats()
    ss = plan.stage_stats()
    for _ in range(args.warmup):
        plan.apply(x, y)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.apply(x, y)
    ctx.sync()
    wall = time.perf_counter() - t0
    us_apply, us1, us3 = plan.time(x, y, max(args.steps, 20))
    ms = 1e3 * wall / args.steps
    dom = "stage1" if us1 >= us3 else "stage3"
    b = ss["s1_bytes"] if dom == "stage1" else ss["s3_bytes"]
    f = ss["s1_flops"] if dom == "stage1" else ss["s3_flops"]
    t_dom = (us1 if dom == "stage1" else us3) * 1e-6
    intensity = f / b if b else 0.0
    ridge = MFMA_PEAK_TFLOPS[args.prec] * 1e12 / (HBM_PEAK_GBS * 1e9)
    if intensity > ridge:
        achieved = f / t_dom / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_PEAK_TFLOPS[args.prec],
                "unit": "TFLOP/s", "frac": round(achieved / MFMA_PEAK_TFLOPS[args.prec], 4)}
    else:
        achieved = b / t_dom / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4)}
    roof.update({"traffic": None, "kernel": dom, "kernel_us": round(t_dom * 1e6, 3), "kernel_bytes": b,
                 "kernel_flops": f, "kernel_tflops": round(f / t_dom / 1e12, 3),
                 "kernel_GBs": round(b / t_dom / 1e9, 1), "flops_per_byte": round(intensity, 2)})
    workload, metric = HEFF_WORKLOADS[args.workload]
    rec = {
        "metric": metric,
        "value": round(1e3 * ms, 3),
        "unit": "us/matvec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3-f32acc"}[args.prec],
        "data": "synthetic",
        "config": dict({"workload": workload, "psi_blocks": st["psi_blocks"], "dim": st["dim"],
                        "precision": args.prec}, **desc),
        "roofline": roof,
        "event_us": {"apply": round(us_apply, 3), "stage1": round(us1, 3), "stage3": round(us3, 3)},
        "work": {"flops": st["flops"], "bytes": st["bytes"], "stage1_items": st["stage1_items"],
                 "stage3_items": st["stage3_items"]},
        "precisions": compare,
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline_heff(p, args.cpu_seconds)
    plan.close()
    ctx.close()
    return rec


def bench_copy(args):
    """Calibration copy of known size through the library's strided copy
    kernel: block_fuse of one 4096 x 4096 FP64 block (128 MiB read, 128 MiB
    written per call).  Counter readings of this run scale the HBM counters of
    the other kernels."""
    import numpy as np

    import tk_hip

    ctx = tk_hip.Context(0)
    legs = [[(0, 4096)], [(0, 4096)]]
    dirs = [tk_hip.DIR_OUT, tk_hip.DIR_IN]
    x = np.arange(4096 * 4096, dtype=np.float64)
    t = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        yv = tk_hip.block_fuse(ctx, tk_hip.SYM_U1, legs, dirs, (0, 0), 0, 2, x)
        t.append(time.perf_counter() - t0)
    if not np.array_equal(yv, x):
        raise SystemExit("bench.py: calibration copy returned wrong data")
    ctx.close()
    return {"metric": "calibration copy (block_fuse of one 4096 x 4096 FP64 block, host round trip)",
            "value": round(1e3 * min(t), 3), "unit": "ms/call", "n_gpus": 1, "steps": args.steps, "warmup": 0,
            "ms_per_step": round(1e3 * min(t), 3), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "copy_4096x4096_f64", "bytes_read": 8 * x.size, "bytes_written": 8 * x.size}}


def run_sweeps(tk_hip, ctx, mpo, sym, bonds, data, sched, warmup, steps, prec, tol, log):
    """The bond-dimension ramp (sweeps at D < D_max, never timed), then
    `warmup` untimed and `steps` timed sweeps at D_max, all from the same
    seeded MPS; kernel timing is on for the timed sweeps only.  Returns every
    sweep's statistics, the timed sweeps' wall times, the kernel record and
    the centre bonds."""
    d = tk_hip.Dmrg(ctx, mpo, sym, prec=prec, schedule=sched, davidson_tol=tol, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    t0 = time.perf_counter()
    d.prepare()
    log(f"[{prec}] prepared in {time.perf_counter() - t0:.2f}s")
    ramp = len(sched) - 1
    first_timed = ramp + warmup
    out, times = [], []
    for k in range(first_timed + steps):
        timed = k >= first_timed
        if timed and k == first_timed:
            ctx.sync()
            ctx.timing(True)
        t0 = time.perf_counter()
        s = d.sweep()
        ctx.sync()
        el = time.perf_counter() - t0
        s["wall"] = el
        s["bond_dim"] = sched[min(k, ramp)]
        if timed:
            times.append(el)
        out.append(s)
        tag = " (timed)" if timed else (" (ramp)" if k < ramp else " (warmup)")
        log(f"[{prec}] sweep {k + 1}/{first_timed + steps}{tag} "
            f"D={s['bond_dim']} E={s['energy']:.13f} t={el:.3f}s (plan {s['t_plan']:.2f} "
            f"eig {s['t_eig']:.2f} svd {s['t_svd']:.2f} env {s['t_env']:.2f}) matvecs={s['matvecs']} "
            f"fp64={s['matvecs_fp64']} maxD={s['max_bond_dim']} trunc={s['max_truncation_error']:.2e}")
    timing = ctx.timing_read()
    ctx.timing(False)
    n = d.n
    centre = (d.bond(n // 2 - 1), d.bond(n // 2 + 1))
    d.close()
    return out, times, timing, centre


# Measured HBM traffic per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated
# with a copy kernel of known size) for kernels of this workload, from the
# summaries in profiles/: "<kernel>": "<file>".
TRAFFIC_PROFILES = {"tridiag_column_step": "tridiag_column_step_traffic_dmrg_sweep_L64_D2048.md"}


def measured_traffic(name):
    f = TRAFFIC_PROFILES.get(name)
    if f is None:
        return None
    try:
        with open(os.path.join(ROOT, "profiles", f)) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        return None
    return None


def roofline_of(timing):
    """Roofline of the kernel with the largest event time: against the HBM
    peak (algorithmic bytes per launch over the mean launch duration), or,
    for a contraction whose algorithmic flops per byte exceed the ridge of
    its precision (the kernel name's suffix), against that MFMA peak (for
    split-BF16 the dense BF16 peak, each product counted once although it
    takes three BF16 MFMAs); traffic is the profiled HBM bytes per launch of
    that kernel when a summary exists."""
    name = max((k for k in timing if not k.startswith("phase_")), key=lambda k: timing[k]["ms"])
    t = timing[name]
    sec = t["ms"] * 1e-3
    common = {"traffic": measured_traffic(name), "kernel": name, "launches": t["launches"],
              "mean_launch_us": round(1e3 * t["ms"] / max(1, t["launches"]), 3),
              "bytes_per_launch": round(t["bytes"] / max(1, t["launches"]), 1),
              "tflops": round(t["flops"] / sec / 1e12, 3) if sec > 0 else 0.0}
    prec = next((p for p in MFMA_PEAK_TFLOPS if name.endswith("_" + p)), None)
    if prec is not None and t["bytes"] > 0 and sec > 0:
        peak = MFMA_PEAK_TFLOPS[prec]
        if t["flops"] / t["bytes"] > peak * 1e12 / (HBM_PEAK_GBS * 1e9):
            achieved = t["flops"] / sec / 1e12
            return dict({"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                         "frac": round(achieved / peak, 4)}, **common)
    achieved = t["bytes"] / sec / 1e9 if sec > 0 else 0.0
    return dict({"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 4)}, **common)


def kernel_table(timing):
    out = {}
    for name, t in sorted(timing.items(), key=lambda kv: -kv[1]["ms"]):
        sec = t["ms"] * 1e-3
        out[name] = {"launches": t["launches"], "ms": round(t["ms"], 3),
                     "GB/s": round(t["bytes"] / sec / 1e9, 1) if sec > 0 else 0.0,
                     "TF/s": round(t["flops"] / sec / 1e12, 3) if sec > 0 else 0.0}
    return out


def cpu_baseline_sweep(tk_hip, sites, log):
    """The oracle's FP64 two-site sweeps (matvecs, SVD, environments; OpenMP
    over the host's cores) on the same chain from the same kind of seeded
    MPS: one sweep at D=64 (untimed ramp), then the timed sweep at D=128,
    the largest bond dimension of the schedule family whose sweep fits the
    CPU budget; and the first D=64 sweep on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    mpo = tk_hip.chain_mpo("hubbard", sites, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (sites // 2, sites // 2), 64, seed=0,
                                     min_sector_dim=0)
    cores = os.cpu_count()
    res = {}
    for nthreads, sched in ((0, (64, 128)), (1, (64,))):
        o = oracle.Dmrg(mpo, sym, schedule=sched, davidson_tol=1e-10, svd_cutoff=1e-12, nthreads=nthreads)
        o.set_state(bonds, data)
        o.prepare()
        for dk in sched:
            t0 = time.perf_counter()
            st = o.sweep()
            res[(nthreads, dk)] = (time.perf_counter() - t0, st)
            log(f"[cpu] oracle sweep D={dk} on {cores if nthreads == 0 else 1} thread(s): "
                f"{res[(nthreads, dk)][0]:.2f}s E={st['energy']:.12f}")
        o.close()
    t128, st128 = res[(0, 128)]
    t64_1, _ = res[(1, 64)]
    return {"value": round(t128, 3), "unit": "s/sweep at D=128", "cores": cores, "kind": "port",
            "sample": f"one oracle FP64 two-site sweep (Davidson tol 1e-10, cutoff 1e-12: matvecs "
                      f"{st128['t_matvec']:.1f}s, SVD {st128['t_svd']:.1f}s, environments {st128['t_env']:.1f}s, "
                      f"the rest Davidson vector work) of the same Hubbard chain L={sites} at D=128 after one at "
                      f"D=64, from the PCG64-seeded MPS, on {cores} OpenMP threads",
            "single_thread": {"value": round(t64_1, 3), "unit": "s/sweep at D=64", "cores": 1,
                              "sample": "the first oracle sweep at D=64 from the same seeded MPS on one thread"},
            "scaled_to_d2048_estimate_s": round(t128 * (2048 / 128) ** 3, 1)}


def tune_report(tk_hip, args):
    """The contraction tilings the autotuner measured (signature -> choice and
    the launch times of the choice and of the default 32 x 128 form)."""
    if args.no_tune:
        return {"tuned": False}
    kinds = {0: "32x32 wavefront", 1: "32x128 workgroup", 2: "64x64 workgroup"}
    t = tk_hip.tune_table()
    return {"tuned": True, "entries": len(t),
            "choices": [{"signature": hex(e["signature"]), "kind": kinds[e["wg_kind"]],
                         "tile_min": e["tile_min"] if e["wg_kind"] == 2 else None,
                         "us_best": round(e["us_best"], 2), "us_default": round(e["us_default"], 2)} for e in t]}


def bench_sweep(args):
    import statistics

    import tk_hip

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    n = args.sites
    sched = tk_hip.BondDimensionSchedule.warmup(args.d0, args.dmax, args.ramp).dims
    mpo = tk_hip.chain_mpo("hubbard", n, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (n // 2, n // 2), sched[0], seed=0,
                                     min_sector_dim=0)
    ctx = tk_hip.Context(0)
    ctx.tune(not args.no_tune)
    res = {}
    for prec in ("fp64", args.mixed):
        res[prec] = run_sweeps(tk_hip, ctx, mpo, sym, bonds, data, sched, args.warmup, args.steps, prec, args.tol,
                               log)
    s64, times64, tim64, centre = res["fp64"]
    slo, timeslo, timlo, _ = res[args.mixed]
    ramp = len(sched) - 1
    timed64, timedlo = s64[ramp + args.warmup:], slo[ramp + args.warmup:]
    assert all(s["bond_dim"] == args.dmax for s in timedlo + timed64)
    med64, medlo = statistics.median(times64), statistics.median(timeslo)
    mean_lo = sum(timeslo) / len(timeslo)
    de = abs(slo[-1]["energy"] - s64[-1]["energy"])
    mv_lo = sum(s["matvecs"] for s in timedlo)
    mv_lo64 = sum(s["matvecs_fp64"] for s in timedlo)
    lp_frac = 1.0 - mv_lo64 / mv_lo if mv_lo else 0.0
    dtype = "f64" if mv_lo64 == mv_lo else f"{args.mixed}/f64"
    rec = {
        "metric": METRIC,
        "value": round(medlo, 4),
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mean_lo, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"dmrg_sweep_hubbard_L{n}_D{args.dmax}", "sites": n, "t": 1.0, "U": 4.0,
                   "filling": "half, Sz=0", "symmetry": "U(1)xU(1)", "schedule": sched,
                   "timed_sweep_bond_dim": args.dmax, "davidson_tol": args.tol, "svd_cutoff": 1e-12,
                   "mixed_precision": args.mixed, "initial_mps": "PCG64 seed 0, binomially concentrated charges",
                   "value_is": f"median wall time of the {args.steps} timed sweeps at D={args.dmax} of the mixed "
                               f"run (ramp sweeps and {args.warmup} warmup sweeps untimed); ms_per_step is their "
                               f"mean"},
        "roofline": roofline_of(timlo),
        "fp64_seconds_per_sweep": round(med64, 4),
        "speedup_mixed_vs_fp64": round(med64 / medlo, 3),
        "first_dmax_sweep_s": {"mixed": round(slo[ramp]["wall"], 4), "fp64": round(s64[ramp]["wall"], 4)},
        "timed_sweeps_mixed": [{"D": s["bond_dim"], "s": round(s["wall"], 4), "matvecs": s["matvecs"],
                                "matvecs_fp64": s["matvecs_fp64"]} for s in timedlo],
        "timed_sweeps_fp64": [{"D": s["bond_dim"], "s": round(s["wall"], 4), "matvecs": s["matvecs"]}
                              for s in timed64],
        "energy_fp64": s64[-1]["energy"],
        "energy_mixed": slo[-1]["energy"],
        "abs_delta_e": de,
        "delta_e_ok": de <= 1e-10,
        "max_bond_dim": slo[-1]["max_bond_dim"],
        "max_truncation_error": slo[-1]["max_truncation_error"],
        "phases_mixed": {k: round(sum(s[k] for s in timedlo) / len(timedlo), 4)
                         for k in ("t_plan", "t_eig", "t_svd", "t_env")},
        "phases_fp64": {k: round(sum(s[k] for s in timed64) / len(timed64), 4)
                        for k in ("t_plan", "t_eig", "t_svd", "t_env")},
        "matvecs_mixed": mv_lo,
        "matvecs_mixed_fp64": mv_lo64,
        "mixed_low_precision_matvec_fraction": round(lp_frac, 4),
        "matvecs_fp64": sum(s["matvecs"] for s in timed64),
        "kernels_mixed": kernel_table(timlo),
        "kernels_fp64": kernel_table(tim64),
        "tiling": tune_report(tk_hip, args),
    }
    ctx.close()
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline_sweep(tk_hip, n, log)
    return rec


def _ground_state(tk_hip, ctx, args, log):
    """Two-site FP64 DMRG ground state of the Hubbard chain (t=1, U=4, half
    filling) ramped to args.dmax; returns (engine, mpo, sym, energy)."""
    n = args.sites
    sched = tk_hip.BondDimensionSchedule.warmup(min(args.d0, args.dmax), args.dmax, args.ramp).dims
    mpo = tk_hip.chain_mpo("hubbard", n, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (n // 2, n // 2), sched[0], seed=0,
                                     min_sector_dim=0)
    d = tk_hip.Dmrg(ctx, mpo, sym, schedule=sched, davidson_tol=args.tol, svd_cutoff=1e-12)
    d.set_state(bonds, data)
    d.prepare()
    e = 0.0
    for k in range(len(sched)):
        t0 = time.perf_counter()
        s = d.sweep()
        e = s["energy"]
        log(f"[ground state] sweep {k + 1}/{len(sched)} D={sched[k]} E={e:.13f} t={time.perf_counter() - t0:.3f}s")
    return d, mpo, sym, e, sched


def _timed_record(metric, value, unit, args, workload, timing, extra):
    rec = {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * value, 3), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": workload,
           "roofline": roofline_of(timing), "kernels": kernel_table(timing)}
    rec.update(extra)
    return rec


def bench_tdvp(args):
    """Real-time two-site TDVP of c+_up psi_0 on the Hubbard chain (the G(t)
    workload of the impurity solver): ground state by two-site DMRG ramped to
    D = dmax, c+_up on site 0, then real-time steps dt = 0.05 in the ancilla
    form at maximum bond dimension dmax (Krylov dimension 30), the last
    args.steps of them timed, once with the reference's fixed Krylov dimension
    and once with the early stop (adaptive_tol 1e-12), both from the same
    state; value is the early-stop step time.  No CPU baseline: the oracle
    has no TDVP."""
    import tk_hip

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = tk_hip.Context(0)
    d, mpo, sym, e0, sched = _ground_state(tk_hip, ctx, args, log)
    n2 = d.apply_site_op(tk_hip.SITE_TYPES["electron"]().ops["Cdagup"], (1, 0))
    phi = ([d.bond(i) for i in range(args.sites + 1)], [d.site(i) for i in range(args.sites)])
    d.close()
    dt, runs = 0.05, {}
    for mode, atol in (("fixed", 0.0), ("adaptive", 1e-12)):
        ket = tk_hip.Dmrg(ctx, tk_hip.complex_time_mpo(mpo), sym, schedule=(args.dmax,), svd_cutoff=1e-12)
        ket.set_state(*tk_hip.complex_state(*phi, "R"))
        ket.prepare()
        wall, steps = 0.0, []
        for k in range(args.warmup + args.steps):
            timed = k >= args.warmup
            if timed and k == args.warmup:
                ctx.sync()
                ctx.timing(True)
            t0 = time.perf_counter()
            st = ket.tdvp_step(-dt, args.dmax, 30, 1e-12, atol)
            ctx.sync()
            el = time.perf_counter() - t0
            if timed:
                wall += el
                steps.append(st)
            log(f"[tdvp {mode}] step {k + 1}/{args.warmup + args.steps}{' (timed)' if timed else ''} t={el:.3f}s "
                f"maxD={st['max_bond_dim']} trunc={st['max_truncation_error']:.2e} matvecs={st['matvecs']} "
                f"log_norm={st['log_norm']:.2e}")
        runs[mode] = (ket, wall / args.steps, steps, ctx.timing_read())
        ctx.timing(False)
    fixed, adapt = runs["fixed"], runs["adaptive"]
    # the two runs' states after the same steps: |<fixed|adaptive>| should be 1
    fidelity = abs(fixed[0].overlap(adapt[0]))
    for r in runs.values():
        r[0].close()
    ctx.close()
    timing, steps = adapt[3], adapt[2]
    workload = {"workload": f"tdvp_real_time_hubbard_L{args.sites}_D{args.dmax}", "sites": args.sites, "t": 1.0,
                "U": 4.0, "filling": "half, Sz=0", "operator": "c+_up on site 0", "dt": dt, "krylov_dim": 30,
                "max_bond_dim": args.dmax, "ground_state_schedule": sched}
    workload["krylov_stop"] = "adaptive 1e-12 (value); fixed dimension 30 as in the reference (fixed_*)"
    return _timed_record(f"TDVP real-time step time, Hubbard L={args.sites}, c+ psi0, D<={args.dmax}",
                         adapt[1], "s/step", args, workload, timing,
                         {"ground_state_energy": e0, "norm2_cdag_psi0": n2,
                          "max_bond_dim": max(s["max_bond_dim"] for s in steps),
                          "max_truncation_error": max(s["max_truncation_error"] for s in steps),
                          "matvecs_per_step": sum(s["matvecs"] for s in steps) / len(steps),
                          "fixed_seconds_per_step": round(fixed[1], 4),
                          "fixed_matvecs_per_step": sum(s["matvecs"] for s in fixed[2]) / len(fixed[2]),
                          "fidelity_fixed_vs_adaptive": fidelity,
                          "phases": {k: round(sum(s[k] for s in steps) / len(steps), 4)
                                     for k in ("t_plan", "t_exp", "t_svd", "t_env")},
                          "cpu_baseline": None})


def bench_sweep1(args):
    """Single-site sweep at the two-site ground state's bond dimensions
    (Hubbard chain ramped to D = dmax by two-site sweeps, then args.warmup
    untimed and args.steps timed single-site sweeps).  No CPU baseline."""
    import tk_hip

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = tk_hip.Context(0)
    d, mpo, sym, e2, sched = _ground_state(tk_hip, ctx, args, log)
    wall, out = 0.0, []
    for k in range(args.warmup + args.steps):
        timed = k >= args.warmup
        if timed and k == args.warmup:
            ctx.sync()
            ctx.timing(True)
        t0 = time.perf_counter()
        s = d.sweep_single_site()
        ctx.sync()
        el = time.perf_counter() - t0
        if timed:
            wall += el
            out.append(s)
        log(f"[single-site] sweep {k + 1}{' (timed)' if timed else ''} E={s['energy']:.13f} t={el:.3f}s "
            f"matvecs={s['matvecs']} maxD={s['max_bond_dim']}")
    timing = ctx.timing_read()
    ctx.timing(False)
    d.close()
    ctx.close()
    workload = {"workload": f"dmrg_single_site_sweep_hubbard_L{args.sites}_D{args.dmax}", "sites": args.sites,
                "t": 1.0, "U": 4.0, "filling": "half, Sz=0", "two_site_schedule": sched}
    return _timed_record(f"single-site DMRG sweep time, Hubbard L={args.sites}, D={args.dmax}",
                         wall / args.steps, "s/sweep", args, workload, timing,
                         {"energy_two_site": e2, "energy_single_site": out[-1]["energy"],
                          "energy_not_raised": out[-1]["energy"] <= e2 + 1e-10,
                          "max_bond_dim": out[-1]["max_bond_dim"], "matvecs": out[-1]["matvecs"],
                          "cpu_baseline": None})


def bench_sweep_c4(args):
    """Config-4 DMRG: J1 Heisenberg on the 6 x 12 cylinder (snake order, MPO
    bond dims up to 20, U(1), Sz = 0), D ramped to 4096, the same seeded MPS
    run in FP64, FP32-MFMA and split-BF16 (both with FP64 refinement); the
    timed step is the sweep at the largest D.  value = split-BF16 sweep time."""
    import tk_hip

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    sched = tk_hip.BondDimensionSchedule.warmup(args.d0, args.dmax, args.ramp).dims
    mpo = tk_hip.cylinder_mpo(6, 12)
    n = len(mpo)
    sym = tk_hip.SYM_U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (0, 0), sched[0], seed=0, min_sector_dim=0)
    ctx = tk_hip.Context(0)
    res = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        res[prec] = run_sweeps(tk_hip, ctx, mpo, sym, bonds, data, sched, args.warmup, args.steps, prec, args.tol,
                               log)
    s64 = res["fp64"][0]
    per = {}
    first = len(sched) - 1 + args.warmup
    for prec, (st, times, tim, _) in res.items():
        per[prec] = {"seconds_per_sweep": round(sum(times) / len(times), 4), "energy": st[-1]["energy"],
                     "abs_delta_e_vs_fp64": abs(st[-1]["energy"] - s64[-1]["energy"]),
                     "max_truncation_error": st[-1]["max_truncation_error"],
                     "max_bond_dim": st[-1]["max_bond_dim"],
                     "matvecs": sum(x["matvecs"] for x in st[first:]),
                     "matvecs_fp64": sum(x["matvecs_fp64"] for x in st[first:])}
    slo, tlos, timlo, _ = res["bf16x3"]
    tlo = sum(tlos) / len(tlos)
    rec = {
        "metric": "DMRG sweep time, J1 Heisenberg 6x12 cylinder D=4096, split-BF16 vs FP32-MFMA vs FP64",
        "value": round(tlo, 4),
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * tlo, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-f32acc (f64 Krylov, SVD, refinement)",
        "data": "synthetic",
        "config": {"workload": "dmrg_sweep_cylinder_6x12_D4096", "sites": n, "J": 1.0, "lattice": "cylinder 6 x 12",
                   "ordering": "snake", "symmetry": "U(1)", "sz": 0, "schedule": sched,
                   "timed_sweep_bond_dim": sched[-1], "davidson_tol": args.tol, "svd_cutoff": 1e-12,
                   "initial_mps": "PCG64 seed 0, binomially concentrated charges"},
        "roofline": roofline_of(timlo),
        "precisions": per,
        "kernels_bf16x3": kernel_table(timlo),
    }
    ctx.close()
    return rec


def bench_permute(args):
    """Block-sparse data movement on the config-4 two-site tensor (legs
    s_i, s_j (spin-1/2), a_L, a_R (the U(1) bond of 4096 states in 31
    sectors), flux 0: 4.66 M FP64 entries): BlockSparseTensor::permute with
    a_L <-> a_R (every block transposed through the LDS-staged tiles), the
    permute (a_L, s_i, s_j, a_R) that keeps the innermost axis (gathered
    rows), and fuse_legs of (s_i, s_j, a_L) (strided chunk copies), each timed
    with HIP events around its kernels on the device-resident tensor (the
    host-built work lists and their uploads excluded); a step is one call.
    value = the transposing permute's algorithmic bandwidth (16 bytes per
    entry).  Each result is checked by the inverse operation reproducing the
    input."""
    import ctypes as C

    import numpy as np
    import tk_hip

    ctx = tk_hip.Context(0)
    sym = tk_hip.SYM_U1
    bond = tk_hip.synthetic_u1_bond(4096)
    legs = [tk_hip.SPIN_PHYS, tk_hip.SPIN_PHYS, bond, bond]
    dirs = [tk_hip.DIR_OUT, tk_hip.DIR_OUT, tk_hip.DIR_OUT, tk_hip.DIR_IN]
    _, _, _, tot = tk_hip.block_table(sym, legs, dirs, (0, 0))
    x = np.random.default_rng(7).standard_normal(tot)
    dx, dy, dz = ctx.to_device(x), ctx.empty(tot), ctx.empty(tot)
    qs, arr, d, f = tk_hip._leg_args(legs, dirs, (0, 0))
    pI = C.POINTER(C.c_int32)

    def permute(p, src, dst, lg, dr):
        q2, a2, d2, f2 = tk_hip._leg_args(lg, dr, (0, 0))
        pp = np.ascontiguousarray(p, dtype=np.int32)
        tk_hip._check(ctx.L.tk_block_permute(ctx.h, sym, len(lg), a2, d2.ctypes.data_as(pI), f2.ctypes.data_as(pI),
                                             pp.ctypes.data_as(pI), C.c_void_p(src.ptr), tot, C.c_void_p(dst.ptr),
                                             tot), ctx.h)

    res = {}
    for name, perm in (("transpose_aL_aR", [0, 1, 3, 2]), ("keep_innermost", [2, 0, 1, 3])):
        for _ in range(args.warmup):
            permute(perm, dx, dy, legs, dirs)
        ctx.timing(True)
        for _ in range(args.steps):
            permute(perm, dx, dy, legs, dirs)
        tim = ctx.timing_read()
        ctx.timing(False)
        inv = [perm.index(k) for k in range(4)]
        permute(inv, dy, dz, [legs[k] for k in perm], [dirs[k] for k in perm])
        ok = bool(np.array_equal(dz.get(), x))
        parts = {k: v for k, v in tim.items() if k in ("copy_tiles", "copy_rows", "permute_gather")}
        us = 1e3 * sum(v["ms"] for v in parts.values()) / args.steps
        res[name] = {"us": round(us, 2), "GB/s": round(16.0 * tot / (us * 1e-6) / 1e9, 1), "round_trip_exact": ok,
                     "kernels": {k: {"us": round(1e3 * v["ms"] / args.steps, 2),
                                     "GB/s": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in parts.items()}}
    for _ in range(args.warmup):
        tk_hip._check(ctx.L.tk_block_fuse(ctx.h, sym, 4, arr, d.ctypes.data_as(pI), f.ctypes.data_as(pI), 0, 3,
                                          C.c_void_p(dx.ptr), tot, C.c_void_p(dy.ptr), tot), ctx.h)
    ctx.timing(True)
    for _ in range(args.steps):
        tk_hip._check(ctx.L.tk_block_fuse(ctx.h, sym, 4, arr, d.ctypes.data_as(pI), f.ctypes.data_as(pI), 0, 3,
                                          C.c_void_p(dx.ptr), tot, C.c_void_p(dy.ptr), tot), ctx.h)
    t = ctx.timing_read()["copy_rows"]
    ctx.timing(False)
    us = 1e3 * t["ms"] / args.steps
    res["fuse_si_sj_aL"] = {"us": round(us, 2), "GB/s": round(t["bytes"] / args.steps / (us * 1e-6) / 1e9, 1)}
    for b in (dx, dy, dz):
        b.free()
    ctx.close()
    if not all(v.get("round_trip_exact", True) for v in res.values()):
        raise SystemExit("bench.py: a permute round trip did not reproduce its input")
    gbs = res["transpose_aL_aR"]["GB/s"]
    return {"metric": "block-sparse permute bandwidth (a_L <-> a_R on the 4.66 M-entry config-4 two-site tensor)",
            "value": gbs, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(res["transpose_aL_aR"]["us"] / 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "block_permute_cylinder_psi_D4096", "entries": int(tot), "blocks": "U(1), 31 sectors"},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": round(gbs / 8000.0, 4),
                         "traffic": None, "kernel": "copy_tiles_kernel"},
            "results": res, "cpu_baseline": None}


def bench_impurity(args):
    """The DMFT impurity workload: args.steps independent impurity problems
    (impurity + 7 bath sites in chain geometry, U = 4, seeded baths), each
    solved by the two-site DMRG ground state at D = 256 and the retarded G(t)
    by real-time two-site TDVP to t = 20 in steps of 0.05 (particle and hole,
    400 steps each), args.workers host threads on their own streams; after
    args.warmup untimed solves.  value = solves per minute of the whole batch;
    the single-solve time (one thread) is reported beside it.  No CPU
    baseline: the oracle has no TDVP (its exact diagonalisation checks G(t)
    in the tests)."""
    import numpy as np
    import tk_dmft

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    rng = np.random.default_rng(2025)
    nb, U = 7, 4.0
    baths = [(np.sort(rng.uniform(-1.5, 1.5, nb)), rng.uniform(0.3, 0.8, nb)) for _ in range(args.steps)]
    kw = {"dt": 0.05, "t_max": 20.0}
    t0 = time.perf_counter()
    warm = tk_dmft.green_time_many(baths[:max(1, args.warmup)], nb, U, workers=1, **kw)
    single = (time.perf_counter() - t0) / max(1, args.warmup)
    log(f"[impurity] single solve (warm-up, one thread): {single:.2f}s; info {warm[0][2]}")
    t0 = time.perf_counter()
    done = []

    def progress(i, info):
        done.append(i)
        log(f"[impurity] problem {i} done ({len(done)}/{len(baths)}) at {time.perf_counter() - t0:.1f}s")

    out = tk_dmft.green_time_many(baths, nb, U, workers=args.workers, progress=progress, **kw)
    wall = time.perf_counter() - t0
    log(f"[impurity] {len(baths)} solves on {args.workers} threads: {wall:.2f}s")
    g0 = np.array([abs(o[1][0] + 1j) for o in out])
    return {"metric": f"impurity G(t) solves per minute (8-site AIM, D=256, TDVP to t=20, dt=0.05), "
                      f"{len(baths)} concurrent",
            "value": round(60.0 * len(baths) / wall, 3), "unit": "solves/min", "n_gpus": 1, "steps": len(baths),
            "warmup": args.warmup, "ms_per_step": round(1e3 * wall / len(baths), 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "dmft_impurity_gt_nbath7_D256_t20", "n_bath": nb, "U": U, "bond_dim": 256,
                       "dt": 0.05, "t_max": 20.0, "problems": len(baths), "threads": args.workers},
            "wall_s": round(wall, 3), "single_solve_s": round(single, 3),
            "speedup_vs_one_thread": round(single * len(baths) / wall, 2),
            "max_abs_G0_plus_i": float(g0.max()), "cpu_baseline": None, "roofline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="sweep",
                    choices=["sweep", "sweep-c4", "heff", "heff-d2048", "heff-c3", "heff-c4", "copy", "tdvp", "sweep1", "impurity", "permute"])
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--mixed", default="bf16x3", choices=["fp32", "bf16x3"])
    ap.add_argument("--sites", type=int, default=64)
    ap.add_argument("--d0", type=int, default=256)
    ap.add_argument("--dmax", type=int, default=None)
    ap.add_argument("--ramp", type=int, default=None)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="run the default contraction tiling (no autotuning)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--workers", type=int, default=16, help="host threads (streams) of the impurity workload")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        args.steps = 3 if args.steps is None else args.steps
        rec = bench_copy(args)
    elif args.workload in HEFF_WORKLOADS:
        args.dmax = 2048 if args.dmax is None else args.dmax
        args.ramp = 7 if args.ramp is None else args.ramp
        args.steps = 200 if args.steps is None else args.steps
        args.warmup = 20 if args.warmup is None else args.warmup
        rec = bench_heff(args)
    elif args.workload in ("tdvp", "sweep1"):
        args.dmax = (256 if args.workload == "tdvp" else 2048) if args.dmax is None else args.dmax
        args.ramp = (4 if args.workload == "tdvp" else 7) if args.ramp is None else args.ramp
        args.steps = (4 if args.workload == "tdvp" else 1) if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        rec = bench_tdvp(args) if args.workload == "tdvp" else bench_sweep1(args)
    elif args.workload == "permute":
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 3 if args.warmup is None else args.warmup
        rec = bench_permute(args)
    elif args.workload == "impurity":
        args.steps = 32 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        rec = bench_impurity(args)
    elif args.workload == "sweep-c4":
        args.dmax = 4096 if args.dmax is None else args.dmax
        args.ramp = 4 if args.ramp is None else args.ramp
        args.steps = 1 if args.steps is None else args.steps
        args.warmup = 0 if args.warmup is None else args.warmup
        rec = bench_sweep_c4(args)
    else:
        # the ramp's sweeps (D < D_max) are never timed; --warmup counts
        # untimed sweeps at D_max after them
        args.dmax = 2048 if args.dmax is None else args.dmax
        args.ramp = 7 if args.ramp is None else args.ramp
        args.steps = 3 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        rec = bench_sweep(args)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
