"""Serial against Jacobi order for the two order-dependent in-place kernels, over coupled steps.

The reference updates the enthalpy in place inside advect_diffuse_temperature (kernels.py:
1664-1682) and phi in place inside apply_contact_angle_phi_bc (kernels.py:952-991), so its result
depends on the order the cells are visited. The GPU path takes every neighbour read from the
state before the kernel (Jacobi form). This script runs the oracle twin on the seeded-pool
bead-on-plate fixture four ways — both kernels in Jacobi form (the GPU's semantics), the thermal
kernel serial, the contact-angle kernel serial, both serial (single-thread lexicographic order,
k fastest) — and prints the relative L2 difference of each run's T, phi, f_l and |u| from the
Jacobi run, the number of cells whose flag differs, and the pool width and depth of each run.

    python tests/measure_serial_jacobi.py --case bench  --steps 100   # 256x128x64 (minutes)
    python tests/measure_serial_jacobi.py --case bead_mrt --steps 100  # 64x40x32 (seconds)

Test infrastructure: it drives the oracle only.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

MODES = ("jacobi", "serial_thermal", "serial_contact", "serial")


def build_oracle(name: str, mode: str, **overrides):
    """The oracle twin of a named step case, reset and seeded as the GPU parity tests seed it."""
    from oracle.oracle import Material, OracleTwin
    from tests.fixtures import LCS_BEAD_ON_PLATE
    from tests.step_cases import case_kwargs, configure, fixture_arrays

    kw = dict(case_kwargs(name), **overrides)
    mat_d = dict(LCS_BEAD_ON_PLATE)
    ot = OracleTwin(Material(**mat_d), mode=mode, **kw)
    configure(ot, name)
    ot.reset()
    og = ot.g
    T, f_l, phi = og.T.copy(), og.f_l.copy(), og.phi.copy()
    flags, H = og.flags.copy(), og.H.copy()
    fixture_arrays(T, f_l, phi, flags, H, ot.nz_solid, mat_d, ot.cp_rho, ot.L_rho)
    for arr_name, arr in (("T", T), ("f_l", f_l), ("phi", phi), ("flags", flags), ("H", H)):
        getattr(og, arr_name)[...] = arr
    return ot


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a, b = a.astype(np.float64), b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run(name: str, steps: int, **overrides) -> dict:
    from tests.step_cases import torch_x

    states = {}
    for mode in MODES:
        ot = build_oracle(name, mode, **overrides)
        g = ot.g
        t0 = time.perf_counter()
        for s in range(steps):
            ot.step(torch_x(g.dx, g.nx, s, g.dt), (g.ny // 2) * g.dx, True)
        speed = np.sqrt(g.ux.astype(np.float64) ** 2 + g.uy ** 2 + g.uz ** 2)
        pm = ot.pool_metrics()
        W, D = pm["pool_width_m"] * 1e3, pm["pool_depth_m"] * 1e3
        states[mode] = dict(T=g.T.copy(), phi=g.phi.copy(), f_l=g.f_l.copy(), u=speed,
                            flags=g.flags.copy(), W=W, D=D)
        print(f"[serial-vs-jacobi] {name} {mode}: {steps} steps in "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ref = states["jacobi"]
    out = {"case": name, "grid": [ref["T"].shape[0], ref["T"].shape[1], ref["T"].shape[2]],
           "steps": steps, "pool_jacobi": {"W_mm": ref["W"], "D_mm": ref["D"]}}
    for mode in MODES[1:]:
        st = states[mode]
        out[mode] = {"rel_l2_vs_jacobi": {k: rel_l2(st[k], ref[k]) for k in ("T", "phi", "f_l", "u")},
                     "flags_differing": int((st["flags"] != ref["flags"]).sum()),
                     "W_mm": st["W"], "D_mm": st["D"]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="bead_mrt", choices=("bead_mrt", "bench"))
    ap.add_argument("--steps", type=int, default=100)
    a = ap.parse_args()
    if a.case == "bench":
        res = run("bead_mrt", a.steps, nx=256, ny=128, nz=64)
        res["case"] = "seeded-pool bead-on-plate 256x128x64"
    else:
        res = run("bead_mrt", a.steps)
        res["case"] = "seeded-pool bead-on-plate 64x40x32"
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
