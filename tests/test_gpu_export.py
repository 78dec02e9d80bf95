"""On-disk formats and telemetry of a GPU run against the oracle.

The export-side kernels (vorticity, max-norm, probe gather) are checked bit for bit against the
oracle or numpy; the exported volume's arrays against the oracle twin's fields after the same
steps, with the reference's unit conversions (reference export/vtk_io.py:85-177); the telemetry's
bead height, toe angle and per-force diagnostics against the oracle's restatements; the bundle,
sequence and probe CSV as the reference's own tests check them (reference
validation/test_export_bundle.py, test_export_full_vtk.py, test_probe_recorder.py,
test_surface_vtk.py).
"""

from __future__ import annotations

import csv
import json

import numpy as np
import pytest

from tests.gpu_common import bits_equal, mismatch_report, random_state
from tests.step_cases import build_pair, run_pair

pytestmark = pytest.mark.gpu

# the reference's telemetry keys (reference twin.py:769-821)
TELEMETRY_KEYS = (
    "step", "sim_time_ms", "peak_temp_K", "peak_temp_C", "peak_cooling_rate_Ks", "pool_width_mm",
    "pool_depth_mm", "pool_length_mm", "n_liquid_cells", "marangoni_vel_ms", "material_status",
    "material_name", "heat_source", "material_tables", "vof_enabled", "travel_speed_mm_s",
    "variable_tau", "porosity_pct", "n_trapped_tracers", "window_offset_x_mm", "deposited_mass_g",
    "expected_wire_mass_g", "expected_drop_mass_g", "n_droplets_fired", "droplet_mass_mg",
    "droplet_transfer_mode", "droplet_impact_velocity_ms", "mass_balance_ratio",
    "wire_mass_flux_g_s", "T_vapor_cap_K", "arc_surface_weighting", "enable_enthalpy_cap",
    "enable_wetting", "enable_hydrostatic_gravity", "enable_bead_freeze", "enable_ctwd",
    "contact_angle_deg", "bead_height_mm", "toe_angle_deg", "ctwd_mm", "T_drop_K",
    "deposition_overflow_count", "physics_tier", "strict_mode", "arc_pressure_model",
    "arc_pressure_peak_pa", "enable_lorentz", "enable_gas_shear", "lorentz_unconverged_count",
    "lorentz_unconverged_streak", "force_diagnostics")


@pytest.fixture(scope="module")
def env(hip_device):
    import torch

    from oracle.oracle import lib as olib
    from waam_solver_amd import kernels as K

    return dict(torch=torch, K=K, O=olib(), dev=hip_device)


def up(env, a):
    return env["torch"].from_numpy(np.ascontiguousarray(a)).to(env["dev"])


def test_vorticity_max_norm_and_probe_gather(env):
    from oracle.oracle import ptr

    dims = (24, 20, 16)
    s = random_state(*dims, seed=31)
    g = {k: up(env, s[k]) for k in ("ux", "uy", "uz", "flags", "T", "f_l", "Fx", "Fy", "Fz")}
    vort = env["torch"].zeros_like(g["ux"])
    o = np.zeros(dims, np.float32)
    dt = 2.0e-5
    env["K"].compute_vorticity_magnitude(g["ux"], g["uy"], g["uz"], g["flags"], vort, 2e-4, dt,
                                         2, *dims)
    env["O"].or_compute_vorticity_magnitude(ptr(s["ux"]), ptr(s["uy"]), ptr(s["uz"]),
                                            ptr(s["flags"]), ptr(o), dt, *dims)
    got = vort.cpu().numpy()
    assert bits_equal(got, o), mismatch_report("vorticity", got, o)
    out = env["torch"].zeros(1, dtype=env["torch"].float32, device=env["dev"])
    env["K"].max_norm3_into(g["Fx"], g["Fy"], g["Fz"], out)
    want = np.sqrt(s["Fx"] * s["Fx"] + s["Fy"] * s["Fy"] + s["Fz"] * s["Fz"]).max()
    assert out.item() == float(want)
    g["Fy"][3, 4, 5] = float("nan")
    env["K"].max_norm3_into(g["Fx"], g["Fy"], g["Fz"], out)
    assert np.isnan(out.item())
    n = int(np.prod(dims))
    cells = np.array([0, 17, n - 1, -1, 4242], np.int64)
    row = env["torch"].zeros(2 * len(cells), dtype=env["torch"].float32, device=env["dev"])
    env["K"].probe_gather_into(g["T"], g["f_l"], up(env, cells), row)
    r = row.cpu().numpy().reshape(-1, 2)
    for (T, fl), c in zip(r, cells):
        if c < 0:
            assert np.isnan(T) and np.isnan(fl)
        else:
            assert T == s["T"].reshape(-1)[c] and fl == s["f_l"].reshape(-1)[c]


def _oracle_derived(ot):
    from oracle.oracle import ptr

    og, L = ot.g, ot.lib
    kap = np.zeros_like(og.phi)
    L.or_compute_curvature_field(ptr(og.phi), ptr(og.flags), ptr(kap), *og.dims)
    vort = np.zeros_like(og.phi)
    L.or_compute_vorticity_magnitude(ptr(og.ux), ptr(og.uy), ptr(og.uz), ptr(og.flags),
                                     ptr(vort), og.dt, *og.dims)
    return kap, vort


@pytest.mark.parametrize("layout", [0, 1])
def test_volume_export_matches_the_oracle_fields(hip_device, tmp_path, layout):
    from waam_solver_amd.export.vtkxml import read_vtk_arrays

    print(f"[export] volume layout={layout}", flush=True)
    tw, ot = build_pair("surface_forces_mrt", layout=layout)
    run_pair(tw, ot, 6)
    out = tw.export_vtk_full(str(tmp_path / "full.vts"), tiers=(0, 1, 2, 3))
    assert out.endswith(".vti")
    attrs, a = read_vtk_arrays(out)
    g, og = tw.grid, ot.g
    assert attrs["extent"] == [0, g.nx, 0, g.ny, 0, g.nz]
    assert attrs["spacing"] == [g.dx * 1000.0] * 3
    assert len(a) >= 15 and "Velocity_Y_ms" in a and "Curvature_kappa" in a
    F = lambda x: np.asarray(x).ravel(order="F")  # noqa: E731  VTK cell order, x fastest
    ms = lambda x: x * g.dx / g.dt  # noqa: E731
    kap, vort = _oracle_derived(ot)
    want = {"Temperature_K": og.T, "Temperature_C": og.T - 273.15, "Liquid_Fraction": og.f_l,
            "VOF_phi": og.phi, "Cell_Flags": og.flags.astype(np.int32), "T_max_K": og.T_max,
            "T_prev_K": og.T_prev, "dTdt_Ks": og.dT_dt, "Cooling_Rate_Ks": -og.dT_dt,
            "Enthalpy_Jm3": og.H, "Time_above_800C_s": og.time_above_800_s,
            "Time_above_1100C_s": og.time_above_1100_s,
            "Time_above_solidus_s": og.time_above_solidus_s, "Velocity_X_ms": ms(og.ux),
            "Velocity_Y_ms": ms(og.uy), "Velocity_Z_ms": ms(og.uz), "Density_lu": og.rho,
            "Pressure_Pa_gauge": (og.rho - 1.0) * (1.0 / 3.0) * tw.mat.rho * (g.dx / g.dt) ** 2,
            "BodyForce_X_ms2": ms(og.Fx) / g.dt, "BodyForce_Y_ms2": ms(og.Fy) / g.dt,
            "BodyForce_Z_ms2": ms(og.Fz) / g.dt, "Curvature_kappa": kap,
            "Vorticity_mag_1s": vort,
            "Mushy_Zone": ((og.f_l > 0.05) & (og.f_l < 0.95)).astype(np.float32)}
    ux, uy, uz = ms(og.ux), ms(og.uy), ms(og.uz)
    want["Speed_ms"] = np.sqrt(ux ** 2 + uy ** 2 + uz ** 2)
    fx, fy, fz = ms(og.Fx) / g.dt, ms(og.Fy) / g.dt, ms(og.Fz) / g.dt
    want["BodyForce_mag_ms2"] = np.sqrt(fx ** 2 + fy ** 2 + fz ** 2)
    for name, w in want.items():
        w = F(w)
        assert a[name].dtype == w.dtype, name
        assert bits_equal(a[name], w), mismatch_report(name, a[name], w)
    # no table material: the material tier adds nothing; the Lorentz arrays only with Lorentz
    assert "RhoCp_Jm3K" not in a and "CurrentDensity_X_Am2" not in a


def test_bundle_sequence_surface_and_telemetry(hip_device, tmp_path):
    print("[export] bundle / sequence", flush=True)
    tw, ot = build_pair("surface_forces_mrt")
    run_pair(tw, ot, 5)
    t = tw.get_telemetry()
    assert set(TELEMETRY_KEYS) <= set(t)
    json.dumps(t)
    assert t["bead_height_mm"] == round(ot.bead_height_m() * 1000.0, 3)
    assert t["toe_angle_deg"] == round(ot.toe_angle_deg(), 1)
    assert t["force_diagnostics"] == ot.force_diagnostics()
    assert t["force_diagnostics"]["f_csf_max"] > 0 and t["force_diagnostics"]["f_recoil_max"] >= 0
    assert t["window_offset_x_mm"] == 0.0 and t["enable_gas_shear"] is True
    # the diagnostics put the live forces back: the next steps still match the oracle
    run_pair(tw, ot, 2, start=5)
    from tests.step_cases import compare

    bad = {k: v for k, v in compare(tw, ot).items() if not v[0]}
    assert not bad, bad

    paths = tw.export_research_bundle(tmp_path / "bundle", tag="t0", tiers=(0, 3))
    assert {"volume", "surface", "telemetry", "meta"} <= set(paths)
    meta = json.loads(open(paths["meta"]).read())
    assert meta["grid"]["nx"] == tw.grid.nx and "unit_conversions" in meta
    assert meta["physics_flags"]["enable_recoil"] is True
    assert meta["telemetry"]["step"] == 7
    from waam_solver_amd.export.vtkxml import read_vtk_arrays

    attrs, s = read_vtk_arrays(paths["surface"])
    assert attrs["Polys"] > 0
    pts = s["Points"].reshape(-1, 3)
    # the free surface sits within a few cells of the plate top
    z_top = tw.nz_solid * tw.grid.dx * 1000.0
    assert np.median(pts[:, 2]) == pytest.approx(z_top, abs=3 * tw.grid.dx * 1000.0)
    np.testing.assert_allclose(s["phi"], 0.5, atol=1e-5)
    vti = tw.export_research_sequence(tmp_path / "seq", n_steps=5, every_n=2, tiers=(0,))
    assert len(vti) == 3
    text = (tmp_path / "seq" / "sequence.pvd").read_text()
    assert 'file="frame_0002/' in text


def test_probe_recorder_like_the_reference(hip_device, tmp_path):
    """Reference validation/test_probe_recorder.py: a probe added later has empty cells in the
    earlier rows; every sample equals the field at its cell."""
    from waam_solver_amd.export.probes import ProbeRecorder

    print("[export] probes", flush=True)
    tw, ot = build_pair("bead_mrt")
    tw.probe_recorder = ProbeRecorder(ring_rows=4)
    tw.probe_recorder.add_grid(4, 5, 3, tw, "sub")
    run_pair(tw, ot, 5)
    tw.probe_recorder.add_grid(6, 5, 4, tw, "torch_1")
    run_pair(tw, ot, 3, start=5)
    path = tmp_path / "probes.csv"
    tw.probe_recorder.write_csv(path)
    rows = list(csv.DictReader(open(path)))
    assert len(rows) == 8
    assert rows[0]["torch_1_T_K"] == "" and rows[-1]["torch_1_T_K"] != ""
    assert [int(r["step"]) for r in rows] == list(range(1, 9))
    last = tw.probe_recorder.rows[-1]
    assert last["sub_T_K"] == round(float(ot.g.T[4, 5, 3]), 2)
    assert last["torch_1_f_l"] == round(float(ot.g.f_l[6, 5, 4]), 4)
