"""Writes dmrg_hubbard_l64.json: the oracle's two-site DMRG on the Hubbard
chain L=64 (t=1, U=4, half filling, Sz=0, U(1)xU(1)) from the seeded MPS of
tk_hip.initial_mps (PCG64 seed 0, D0=64), bond dimensions 64 -> 128 -> 256,
Davidson tol 1e-12 (the workload runs 1e-10), SVD cutoff 1e-12: per sweep the energy, the largest bond
dimension and truncation error and every bond's QIndex.  Also the centre bond
QIndices of the same chain at D=2048, read from a GPU run of that workload
(the H_eff problem of the full-size parity test is built on them).

Run from the repository root: python tests/golden/make_dmrg_hubbard_l64.py
(about three minutes on 8 cores).
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "tensorkraft-design_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
import tk_hip  # noqa: E402

N = 64
SCHEDULE = (64, 128, 256)
# tighter than the workload's 1e-10 so that the two eigensolvers' different
# trajectories agree to 1e-10 in the energy of the heavily truncated first sweeps
DAVIDSON_TOL = 1e-12


def main():
    mpo = tk_hip.chain_mpo("hubbard", N, t=1.0, U=4.0)
    sym = tk_hip.SYM_U1U1
    bonds, data = tk_hip.initial_mps(sym, [m.phys for m in mpo], (N // 2, N // 2), SCHEDULE[0], seed=0,
                                     min_sector_dim=0)
    o = oracle.Dmrg(mpo, sym, schedule=SCHEDULE, davidson_tol=DAVIDSON_TOL, svd_cutoff=1e-12)
    o.set_state(bonds, data)
    o.prepare()
    sweeps = []
    for k in range(len(SCHEDULE)):
        t0 = time.time()
        s = o.sweep()
        sweeps.append({"bond_dim": SCHEDULE[k], "energy": s["energy"],
                       "max_bond_dim": int(s["max_bond_dim"]),
                       "max_truncation_error": s["max_truncation_error"],
                       "bonds": [[[list(q), d] for q, d in o.bond(i)] for i in range(N + 1)]})
        print(f"sweep {k + 1} D={SCHEDULE[k]} E={s['energy']:.15f} {time.time() - t0:.1f}s", flush=True)
    o.close()
    out = {"sites": N, "t": 1.0, "U": 4.0, "total": [N // 2, N // 2], "seed": 0, "schedule": list(SCHEDULE),
           "davidson_tol": DAVIDSON_TOL, "svd_cutoff": 1e-12, "sweeps": sweeps,
           "centre_d2048": CENTRE_D2048}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dmrg_hubbard_l64.json"), "w") as f:
        json.dump(out, f, indent=0)


# Bonds 31 and 33 of the D=2048 state after the ramp 256 -> 2048 of the same
# chain (one GPU run of the default benchmark workload).
CENTRE_D2048 = {
    "bond_l": [[[12, 16], 1], [[12, 17], 2], [[12, 18], 1], [[13, 14], 1], [[13, 15], 10], [[13, 16], 27], [[13, 17], 28], [[13, 18], 13], [[13, 19], 1], [[14, 13], 1], [[14, 14], 19], [[14, 15], 74], [[14, 16], 115], [[14, 17], 89], [[14, 18], 28], [[14, 19], 2], [[15, 13], 10], [[15, 14], 74], [[15, 15], 178], [[15, 16], 207], [[15, 17], 115], [[15, 18], 27], [[15, 19], 1], [[16, 12], 1], [[16, 13], 27], [[16, 14], 115], [[16, 15], 207], [[16, 16], 178], [[16, 17], 74], [[16, 18], 10], [[17, 12], 2], [[17, 13], 28], [[17, 14], 89], [[17, 15], 115], [[17, 16], 74], [[17, 17], 19], [[17, 18], 1], [[18, 12], 1], [[18, 13], 13], [[18, 14], 28], [[18, 15], 27], [[18, 16], 10], [[18, 17], 1], [[19, 13], 1], [[19, 14], 2], [[19, 15], 1]],
    "bond_r": [[[13, 17], 1], [[13, 18], 2], [[13, 19], 1], [[14, 15], 1], [[14, 16], 10], [[14, 17], 27], [[14, 18], 28], [[14, 19], 13], [[14, 20], 1], [[15, 14], 1], [[15, 15], 19], [[15, 16], 74], [[15, 17], 115], [[15, 18], 89], [[15, 19], 28], [[15, 20], 2], [[16, 14], 10], [[16, 15], 74], [[16, 16], 178], [[16, 17], 207], [[16, 18], 115], [[16, 19], 27], [[16, 20], 1], [[17, 13], 1], [[17, 14], 27], [[17, 15], 115], [[17, 16], 207], [[17, 17], 178], [[17, 18], 74], [[17, 19], 10], [[18, 13], 2], [[18, 14], 28], [[18, 15], 89], [[18, 16], 115], [[18, 17], 74], [[18, 18], 19], [[18, 19], 1], [[19, 13], 1], [[19, 14], 13], [[19, 15], 28], [[19, 16], 27], [[19, 17], 10], [[19, 18], 1], [[20, 14], 1], [[20, 15], 2], [[20, 16], 1]],
}

if __name__ == "__main__":
    main()
