"""Full-size and full-length runs on the GPU.

- The 4000-step bead-on-plate run on 256x128x64 cells (the settings, seeded noise and torch path
  of bench.py's ``bead_on_plate`` workload, with a molten dome seeded under the torch's start so
  that a pool exists within the run): pool width and depth every 250 steps against the oracle's
  run of the same workload (tests/golden/c2_bead_on_plate_telemetry.json, written by
  tests/golden/make_c2_telemetry.py), within one cell, and the liquid cell count, peak
  temperature and droplet ledger equal.
- The 768x384x192 step with binary16 population storage: 64 steps as two captured graphs against
  64 direct steps from the same state, every field and both population buffers equal; fields
  finite; the binary64 sum of the enthalpy (J) gains the arc's energy over those steps, and every
  metal cell holds rho = 1 up to the collision's mass defect.
- The 256x128x64 workload with droplets, 400 steps with and without row-segment skipping (quiet
  gas, cold plate, the plate march): every field bit-equal.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "c2_bead_on_plate_telemetry.json")


def test_c2_pool_telemetry_over_4000_steps_matches_the_oracle(hip_device):
    import bench

    with open(GOLDEN) as fh:
        ref = json.load(fh)
    from tests.golden.make_c2_telemetry import seed_state
    from waam_solver_amd.twin import WAAMTwin

    w = bench.WORKLOAD
    print(f"[c2] {w['nx']}x{w['ny']}x{w['nz']} x {ref['steps']} steps", flush=True)
    tw = WAAMTwin(material="LCS-bead-on-plate", **bench.twin_kwargs(w))
    bench.configure(tw, w)
    tw.reset()
    g = tw.grid
    arrs = {n: getattr(g, n).to_numpy() for n in ("T", "f_l", "phi", "flags", "H")}
    seed_state(arrs["T"], arrs["f_l"], arrs["phi"], arrs["flags"], arrs["H"], tw.nz_solid,
               tw.mat.T_liquidus, tw.cp_rho, tw.L_rho, w)
    for n, a in arrs.items():
        getattr(g, n).from_numpy(a)
    rows = {r["step"]: r for r in ref["rows"]}
    seen = 0
    for s in range(ref["steps"]):
        tw.step(*bench.torch_xy(w, s, g.dt), True)
        r = rows.get(s + 1)
        if r is None:
            continue
        tel = tw.get_telemetry()
        seen += 1
        W, D = tel["pool_width_mm"] * 1e-3, tel["pool_depth_mm"] * 1e-3
        print(f"[c2] step {s + 1}: W {W * 1e3:.3f} / {r['pool_width_m'] * 1e3:.3f} mm, "
              f"D {D * 1e3:.3f} / {r['pool_depth_m'] * 1e3:.3f} mm, "
              f"liquid {tel['n_liquid_cells']} / {r['n_liquid_cells']}", flush=True)
        assert abs(W - r["pool_width_m"]) <= g.dx * (1 + 1e-9), s + 1
        assert abs(D - r["pool_depth_m"]) <= g.dx * (1 + 1e-9), s + 1
        assert tel["n_liquid_cells"] == r["n_liquid_cells"], s + 1
        assert abs(tel["peak_temp_K"] - round(r["T_peak"], 1)) <= 0.05 + 1e-9, s + 1
        assert tw._n_droplets_fired == r["n_droplets"], s + 1
        assert tw._deposited_volume_m3 == r["deposited_volume_m3"], s + 1
    assert seen == len(ref["rows"])
    # the run melts: the last rows carry a pool
    assert ref["rows"][-1]["pool_width_m"] > 0.0


def test_c5_captured_steps_equal_direct_steps_at_full_size(hip_device):
    import torch

    import bench
    from waam_solver_amd.field import tensor_of

    w = bench.LARGE
    n = 64
    print(f"[c5] {w['nx']}x{w['ny']}x{w['nz']} fp16, {n} captured vs {n} direct", flush=True)
    a = bench.build_gpu_twin(w, "fp16")
    b = bench.build_gpu_twin(w, "fp16")
    ga, gb = a.grid, b.grid

    def ledger(t):
        t.sync()
        g = t.grid
        metal = tensor_of(g.flags) != 2
        return (float(tensor_of(g.H).double().sum().item()) * g.dx ** 3,
                float(tensor_of(g.rho).double()[metal].sum().item()), int(metal.sum().item()))

    xs = [bench.torch_xy(w, s, ga.dt) for s in range(n + 1)]
    # one direct step each (it forms the contact-angle film on the bare plate and lifts the
    # seeded noise of the exposed cells to the enthalpy floor), then n steps: direct on a, as
    # captured graphs of 32 on b; the ledger is kept on a
    a.step(*xs[0], True)
    b.step(*xs[0], True)
    H1, r1, m1 = ledger(a)
    for x, y in xs[1:]:
        a.step(x, y, True)
    b._fused.graph_steps = 32
    b.step_many([p[0] for p in xs[1:]], [p[1] for p in xs[1:]], True)
    torch.cuda.synchronize()
    assert b._fused.captured_steps == n
    a.sync()
    b.sync()
    from tests.step_cases import FIELDS

    diff = [f for f in FIELDS + ("f_src", "f_dst")
            if not torch.equal(tensor_of(getattr(ga, f)), tensor_of(getattr(gb, f)))]
    assert not diff, f"captured steps differ from direct ones in {diff}"
    for f in ("H", "T", "f_l", "phi", "rho", "ux", "uy", "uz", "f_src"):
        assert bool(torch.isfinite(tensor_of(getattr(gb, f))).all().item()), f
    H2, r2, m2 = ledger(a)
    e_arc = w["arc_efficiency"] * w["arc_power_W"] * n * gb.dt
    print(f"[c5] after step 1: sum H {H1:.4f} J, sum rho(metal) {r1:.2f} over {m1} cells; after "
          f"{n} more: gain {H2 - H1:.5f} J (arc {e_arc:.5f} J), sum rho(metal) {r2:.2f} over "
          f"{m2} cells", flush=True)
    # past the first step the enthalpy gains the arc's energy (no cell is near the ceiling; the
    # losses of a plate at ambient temperature are below 1e-3 of it)
    assert abs((H2 - H1) - e_arc) <= 1e-2 * e_arc
    # every metal cell holds rho = 1 up to the collision's mass defect
    assert abs(r2 - m2) <= 1e-4 * m2 and abs(r1 - m1) <= 1e-4 * m1


def test_segment_skipping_is_bit_equal_over_400_steps(hip_device):
    """The 256x128x64 bead-on-plate workload (seeded molten dome, droplets every 100 steps,
    binary16 populations) stepped 400 times by two twins, one skipping quiet gas segments and
    running the plate's cold segments by the plate march (the flag-update and stream sweeps
    left out), one sweeping every cell: every field and both population buffers bit-equal at
    the end, with droplets placed and cold segments left out along the way."""
    import bench
    from tests.golden.make_c2_telemetry import seed_state
    from tests.step_cases import FIELDS, segment_record
    from waam_solver_amd.solvers.fused import FusedStepper
    from waam_solver_amd.twin import WAAMTwin

    w = bench.WORKLOAD
    n = 400
    print(f"[skip] {w['nx']}x{w['ny']}x{w['nz']} fp16, {n} steps skipping vs sweeping", flush=True)
    kw = dict(bench.twin_kwargs(w), droplet_freq_hz=500.0)
    twins = []
    for _ in range(2):
        tw = WAAMTwin(material="LCS-bead-on-plate", population_storage="fp16", **kw)
        bench.configure(tw, w)
        tw.reset()
        g = tw.grid
        arrs = {k: getattr(g, k).to_numpy() for k in ("T", "f_l", "phi", "flags", "H")}
        seed_state(arrs["T"], arrs["f_l"], arrs["phi"], arrs["flags"], arrs["H"], tw.nz_solid,
                   tw.mat.T_liquidus, tw.cp_rho, tw.L_rho, w)
        for k, a in arrs.items():
            getattr(g, k).from_numpy(a)
        twins.append(tw)
    skip, ref = twins
    ref._fused = FusedStepper(ref)
    ref._fused.quiet_skip = False
    g = skip.grid
    cold = 0
    for s in range(n):
        xy = bench.torch_xy(w, s, g.dt)
        skip.step(*xy, True)
        ref.step(*xy, True)
        if s % 50 == 49:
            cold = max(cold, segment_record(skip)[2])
    skip.sync()
    ref.sync()
    diff = [f for f in FIELDS + ("f_src", "f_dst")
            if getattr(skip.grid, f).to_numpy().tobytes() != getattr(ref.grid, f).to_numpy().tobytes()]
    print(f"[skip] droplets {skip._n_droplets_fired}, cold segments left out (max) {cold}, "
          f"differing {diff}", flush=True)
    assert not diff, f"segment skipping changed {diff}"
    assert skip._n_droplets_fired == ref._n_droplets_fired >= 2
    assert cold > 0
