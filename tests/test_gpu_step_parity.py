"""Coupled-step parity: the GPU twin against the oracle twin over 100 steps.

The bar: fp32 fields within relative L2 1e-5 and cell flags identical after 100 steps. Both
sides run the same fixture along the same torch path; because every kernel is evaluated with
the same operand order and rounding, the fields are expected to agree bit for bit, which is
asserted too and reported per field.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.step_cases import (FIELDS, build_pair, cold_skipped_segments, compare, run_pair,
                              segment_record)

TABLE_FIELDS = ("cp_rho_field", "alpha_lu_field", "dgamma_lu_field", "tau_field")

pytestmark = pytest.mark.gpu

REL_L2_TOL = 1e-5


def _assert_parity(res: dict, label: str) -> None:
    bad = {k: v for k, v in res.items() if not v[0]}
    print(f"[{label}] " + ", ".join(f"{k}:{'=' if v[0] else f'{v[1]:.2e}'}" for k, v in res.items()))
    assert res["flags"][0], f"{label}: flags differ"
    for k, (eq, l2) in res.items():
        assert l2 <= REL_L2_TOL, f"{label}: {k} rel L2 {l2:.3e} > {REL_L2_TOL}"
    assert not bad, f"{label}: not bit-identical: {sorted(bad)}"


@pytest.mark.parametrize("engine,layout", [("reference", 1), ("fused", 1), ("fused", 0),
                                           ("reference", 0)])
@pytest.mark.parametrize("case", ["bead_mrt", "droplets_srt", "ctwd_srt", "surface_forces_mrt",
                                  "recoil_const_srt", "full_tier_mrt", "bead_box_mrt",
                                  "bead_odd_mrt"])
def test_coupled_step_parity(hip_device, case, engine, layout):
    """layout 1: fields stored (nz, ny, nx), i fastest; layout 0: the reference's (nx, ny, nz).
    The fused engine in layout 0 runs the surface list in a launch of its own (the large-grid
    choice), in layout 1 in the thermal sweep's launch."""
    tw, ot = build_pair(case, engine=engine, layout=layout)
    if engine == "fused" and layout == 0:
        from waam_solver_amd.solvers.fused import FusedStepper

        tw._fused = FusedStepper(tw)
        tw._fused.shared_thermal_max_cells = 0
    done = 0
    for n in (1, 9, 40, 50):
        run_pair(tw, ot, n, start=done)
        done += n
        _assert_parity(compare(tw, ot), f"{case}/{engine}/layout{layout} step {done}")
    if case in ("droplets_srt", "ctwd_srt", "recoil_const_srt"):
        assert tw._n_droplets_fired == ot._n_droplets_fired and tw._n_droplets_fired > 0
        assert tw._deposited_volume_m3 == ot._deposited_volume_m3
    if case == "ctwd_srt":
        assert tw._bead_height_m == ot._bead_height_m and tw.ctwd_m == ot.ctwd_m
    if engine == "fused" and case in ("bead_mrt", "bead_box_mrt", "bead_odd_mrt"):
        listed, skipped = cold_skipped_segments(tw)
        print(f"[{case}] last steady step: {listed} listed segments, {skipped} cold", flush=True)
        if layout == 1:
            # steady steps left cold plate segments (rows along x inside the plate) out of the
            # flag-update and stream sweeps; in layout 0 a segment runs along z, through the
            # plate's top, and is never all SOLID on these grids
            assert skipped > 0
    if case == "bead_box_mrt" and engine == "fused":
        # the arc's zero-heat box (half-width sqrt(220) sigma + 2 cells) is narrower than the
        # grid in x, so the box branch of the arc sums and the box gate ran
        p = tw._fused._params(2.0, 3.0, True, None)
        assert 2.0 * (220.0 ** 0.5 * p.sigma + 2.0) + 1.0 < tw.grid.nx
    tel = tw.get_telemetry()
    pm = ot.pool_metrics()
    assert tel["n_liquid_cells"] == pm["n_liquid_cells"]
    assert abs(tel["pool_width_mm"] - round(pm["pool_width_m"] * 1000, 3)) < 1e-9
    assert abs(tel["pool_depth_mm"] - round(pm["pool_depth_m"] * 1000, 3)) < 1e-9


@pytest.mark.parametrize("engine", ["fused", "reference"])
@pytest.mark.parametrize("layout", [1, 0])
@pytest.mark.parametrize("first_welding", [True, False])
def test_table_material_parity(hip_device, layout, first_welding, engine):
    """Temperature-dependent cp, k, mu and dgamma/dT (per-cell properties, variable-tau SRT),
    through the fused step (properties from the tables inside its sweeps) and the
    reference-order kernels.

    The per-cell property fields start at zero and are first written by the refresh that
    precedes the thermal update, so a first step that welds clamps enthalpy against
    cp_rho = 0 (reference solvers/coupled_step.py:104-108 before :168-182); the other
    variant starts with one step without the arc so the clamp sees refreshed properties."""
    print(f"[tables] engine {engine} layout {layout} first_welding {first_welding}", flush=True)
    tw, ot = build_pair("tables_mrt", layout=layout, engine=engine)
    g = tw.grid
    done = 0
    if not first_welding:
        x, y = (g.nx // 3) * g.dx, (g.ny // 2) * g.dx
        tw.step(x, y, False)
        ot.step(x, y, False)
    for n in (1, 9, 40, 50):
        run_pair(tw, ot, n, start=done)
        done += n
        _assert_parity(compare(tw, ot, FIELDS + TABLE_FIELDS),
                       f"tables/layout{layout}/weld0={first_welding} step {done}")
    # the refreshed properties span the table range where the metal is hot
    cp = tw.grid.cp_rho_field.to_numpy()
    assert float(cp.max()) > float(cp.min())
    if engine == "fused":
        assert tw._fused is not None  # the fused stepper ran the steps


@pytest.mark.parametrize("engine", ["fused", "reference"])
def test_job_drop_in_parity(hip_device, tmp_path, engine):
    """``WAAMTwin.from_job`` + ``run_path`` on the bead-on-plate job (table material with the
    Sahoo dgamma/dT table, Goldak source, Lin-Eagar pressure, droplets) on a reduced domain,
    replayed step for step on an oracle twin with the same settings."""
    import yaml

    from tests.step_cases import oracle_like
    from waam_solver_amd.job import PKG_ROOT
    from waam_solver_amd.twin import WAAMTwin

    job = yaml.safe_load(open(PKG_ROOT / "jobs/examples/bead_on_plate.yaml"))
    job["simulation"]["domain_mm"] = [16, 10, 8]
    job["plate"] = {"thickness_mm": 3.0, "size_mm": [12, 8]}
    job.pop("torch_path_csv")
    job["torch_path"] = [{"x_mm": 4.0, "y_mm": 5.0}, {"x_mm": 12.0, "y_mm": 5.0}]
    p = tmp_path / "job.yaml"
    p.write_text(yaml.safe_dump(job))
    tw = WAAMTwin.from_job(p)
    # the job's droplet length sets ~36 Hz; detach every ~16 steps instead so that 60 steps
    # deposit
    tw.droplet_freq = 2500.0
    tw.engine = engine
    tw.reset()
    ot = oracle_like(tw)
    ot.reset()
    calls = []
    step_many = tw.step_many

    # run_path hands its positions to step_many (runs of captured steps, droplet steps
    # included, and direct steps); record them there
    def recording_step_many(xs, ys, is_welding=True, zs=None):
        for k, (x, y) in enumerate(zip(xs, ys)):
            calls.append((x, y, is_welding, None if zs is None else zs[k]))
        step_many(xs, ys, is_welding, zs)

    tw.step_many = recording_step_many
    tw.run_path(job, n_steps=60)
    for x, y, w, z in calls:
        ot.step(x, y, w, z)
    assert len(calls) == 60
    if engine == "fused":
        assert tw._fused.captured_steps > 0
    _assert_parity(compare(tw, ot, FIELDS + TABLE_FIELDS), f"job/{engine} step 60")
    assert tw._n_droplets_fired == ot._n_droplets_fired and tw._n_droplets_fired > 0


@pytest.mark.parametrize("engine,layout", [("fused", 1), ("reference", 0)])
def test_moving_window_parity(hip_device, engine, layout):
    """Three window shifts in 100 steps (torch at 0.2 cells per step from 18 cells; trigger at
    0.62 x 32 cells, shifts of 8): every field, both population buffers, the torch offset and
    the tracers follow the oracle."""
    tw, ot = build_pair("window_mrt", engine=engine, layout=layout)
    g = tw.grid
    y = (g.ny // 2) * g.dx
    for s in range(100):
        x = (18.0 + 0.2 * s) * g.dx
        tw.step(x, y, True)
        ot.step(x, y, True)
        if s in (0, 12, 55, 99):
            _assert_parity(compare(tw, ot), f"window/{engine}/layout{layout} step {s + 1}")
    assert tw._window_offset_x_m == ot._window_offset_x_m == 3 * 8 * g.dx


@pytest.mark.parametrize("case", ["bench_small", "droplets_srt", "tables_mrt"])
def test_quiet_segments_skip_with_identical_results(hip_device, case):
    """Two fused twins from the same state, one skipping the sweeps on quiet row segments (all
    GAS with phi 0 around them for two steps) and the flag-update and stream sweeps on cold
    plate segments, one sweeping every cell: every field and both population buffers stay
    identical over 30 steps (with droplets, whose deposition writes cells between steps), and
    the skipping twin did skip segments of both kinds."""
    import torch

    from waam_solver_amd.field import tensor_of
    from waam_solver_amd.solvers.fused import FusedStepper

    tw, _ = build_pair(case)
    ref, _ = build_pair(case)
    ref._fused = FusedStepper(ref)
    ref._fused.quiet_skip = False
    g = tw.grid
    skipped, cold = [], []
    for s in range(30):
        x = (g.nx // 3) * g.dx + 8e-3 * s * g.dt
        y = (g.ny // 2) * g.dx
        tw.step(x, y, True)
        ref.step(x, y, True)
        steady = tw._fused.steady() and s > 0
        nseg, listed, nc = segment_record(tw)
        skipped.append(nseg - listed if steady else 0)
        cold.append(nc if steady else 0)
    torch.cuda.synchronize()
    tw.sync()
    ref.sync()
    diff = [n for n in FIELDS + ("f_src", "f_dst")
            if getattr(tw.grid, n).to_numpy().tobytes() != getattr(ref.grid, n).to_numpy().tobytes()]
    print(f"[quiet] {case}: quiet segments skipped per step {skipped}, cold {cold}")
    assert not diff, f"{case}: quiet skipping changed {diff}"
    assert max(skipped) > 0
    if g.layout == 1:
        assert max(cold) > 0


@pytest.mark.parametrize("pop,nx", [("fp32", 64), ("fp32", 63), ("fp16", 64), ("fp16", 62),
                                    ("fp16", 63), ("bf16", 64)])
def test_pull_bounces_back_for_a_cell_that_turned_gas(hip_device, pop, nx):
    """A FLUID cell in the gas under a SOLID lump turns GAS on the step after its collision:
    the deferred pull of that collision's stream must bounce its downward population back off
    the lump, as the reference's stream does with the flags of the step it belongs to (the cell
    was FLUID then), whatever its flag now. Every pull variant: 8-byte words (nx 64), cell pairs
    (16-bit storage, nx 62), cell by cell (nx 63)."""
    print(f"[pull] turned-gas bounce-back {pop} nx={nx}", flush=True)
    from tests.fixtures import FLUID, SOLID

    tw, ot = build_pair("bead_mrt", population_storage=pop, nx=nx)
    g, og = tw.grid, ot.g
    k0, i0, j0 = tw.nz_solid + 6, nx // 2 + 4, 8
    cp_rho, L_rho = ot.cp_rho, ot.L_rho
    cells = {(i0, j0, k0 + 1): (SOLID, 1.0, 0.0, 1000.0, cp_rho * 1000.0),
             (i0, j0, k0): (FLUID, 0.6, 1.0, 2100.0, cp_rho * 2100.0 + L_rho)}
    arrs = {n: getattr(og, n).copy() for n in ("flags", "phi", "f_l", "T", "H")}
    for (i, j, k), (fl, ph, liq, T, H) in cells.items():
        arrs["flags"][i, j, k] = fl
        arrs["phi"][i, j, k] = ph
        arrs["f_l"][i, j, k] = liq
        arrs["T"][i, j, k] = T
        arrs["H"][i, j, k] = np.float32(H)
    for n, a in arrs.items():
        getattr(og, n)[...] = a
        getattr(g, n).from_numpy(a)
    run_pair(tw, ot, 1)
    assert int(g.flags[i0, j0, k0]) == FLUID and int(og.flags[i0, j0, k0]) == FLUID
    g.phi[i0, j0, k0] = 0.0
    og.phi[i0, j0, k0] = 0.0
    run_pair(tw, ot, 1, start=1)
    assert int(og.flags[i0, j0, k0]) == 2 and int(og.flags[i0, j0, k0 + 1]) == SOLID
    _assert_parity(compare(tw, ot), f"pull/{pop}/nx{nx}")
