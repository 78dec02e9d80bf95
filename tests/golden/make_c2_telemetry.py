"""Writes c2_bead_on_plate_telemetry.json: the oracle's pool telemetry along the 4000-step
bead-on-plate run on 256x128x64 cells (the settings, seeded noise and torch path of bench.py's
``bead_on_plate`` workload) with a molten dome of radius 8 cells seeded on the plate under the
torch's start (the arc alone takes far longer than 4000 steps to melt the plate: the peak
temperature reaches 807 K by step 2250), every 250 steps.

The oracle runs in its Jacobi mode on OpenMP threads; every per-cell loop writes only its own
cell from pre-loop values and the arc weight sum is serial, so the result does not depend on the
thread count. Run from the repository root: ``python tests/golden/make_c2_telemetry.py``.
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "c2_bead_on_plate_telemetry.json")
STEPS = 4000
EVERY = 250


def seed_state(T, f_l, phi, flags, H, nz_solid, T_liquidus, cp_rho, L_rho, w) -> None:
    """The workload's seeded noise, then the molten dome under the torch's start (in place)."""
    import bench
    from tests.fixtures import seed_pool

    bench.seed_noise(T, H, flags, cp_rho)
    seed_pool(T, f_l, phi, flags, H, nz_solid, T_liquidus, np.float32(cp_rho),
              np.float32(L_rho), radius=8,
              center=(int(round(w["x0_mm"] * 1e-3 / w["dx"])), w["ny"] // 2))


def main() -> int:
    import bench
    from oracle.oracle import Material, OracleTwin
    from waam_solver_amd.materials import load_material

    w = bench.WORKLOAD
    m = load_material("LCS-bead-on-plate")
    mat = Material(name=m.name, k=m.k, rho=m.rho, cp=m.cp, T_solidus=m.T_solidus,
                   T_liquidus=m.T_liquidus, L_fusion=m.L_fusion, gamma_0=m.gamma_0,
                   dgamma_dT=m.dgamma_dT, mu=m.mu, beta_T=m.beta_T, alpha=m.alpha)
    ot = OracleTwin(mat, parallel=True, mode="jacobi", **bench.twin_kwargs(w))
    bench.configure(ot, w)
    ot.reset()
    g = ot.g
    seed_state(g.T, g.f_l, g.phi, g.flags, g.H, ot.nz_solid, mat.T_liquidus, ot.cp_rho, ot.L_rho,
               w)
    rows = []
    t0 = time.perf_counter()
    for s in range(STEPS):
        ot.step(*bench.torch_xy(w, s, ot.g.dt), True)
        if (s + 1) % EVERY == 0:
            pm = ot.pool_metrics()
            rows.append({"step": s + 1, "n_liquid_cells": pm["n_liquid_cells"],
                         "pool_width_m": pm["pool_width_m"], "pool_depth_m": pm["pool_depth_m"],
                         "T_peak": pm["T_peak"], "n_droplets": ot._n_droplets_fired,
                         "deposited_volume_m3": ot._deposited_volume_m3})
            print(f"[c2] step {s + 1}: {rows[-1]} ({time.perf_counter() - t0:.0f} s)",
                  flush=True)
    with open(OUT, "w") as fh:
        json.dump({"workload": "bench.py bead_on_plate (256x128x64, dx 0.2 mm, seed 1234) with a "
                               "molten dome of radius 8 under the torch's start",
                   "grid": [w["nx"], w["ny"], w["nz"]], "dx_m": w["dx"], "steps": STEPS,
                   "every": EVERY, "rows": rows}, fh, indent=1)
    print(f"[c2] wrote {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
