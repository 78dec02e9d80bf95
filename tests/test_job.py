"""Job files, material resolution and calibration overlays (host logic, no kernels run).

The twin is built on the CPU device here: construction, job parsing and every host-side
derived constant run without a GPU; no kernel is launched.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from tests.step_cases import derived_constants, oracle_like

BEAD_JOB = "jobs/examples/bead_on_plate.yaml"


def test_material_resolution_order():
    from waam_solver_amd.materials import load_material, resolve_material_path

    assert resolve_material_path("ER70S-6").parent.name == "materials"
    assert resolve_material_path("ER70S-6.v1").parent.name == "validated"
    p = resolve_material_path("materials/validated/SS316L.v1.yaml")
    assert p is not None and p.name == "SS316L.v1.yaml"
    assert resolve_material_path("no-such-alloy") is None
    with pytest.raises(KeyError):
        load_material("no-such-alloy")
    assert not load_material("ER70S-6").tables.has_any()
    assert load_material("SS316L.v1").tables.has_any()


def test_sahoo_table_is_the_temperature_derivative_of_gamma():
    """The Sahoo/DebRoy dgamma/dT knots against a central difference of gamma(T) from the same
    isotherm (Sahoo, DebRoy & McNallan 1988): a wrong sign or a dropped term in either fails."""
    from waam_solver_amd.materials import dgamma_dT_sahoo, gamma_sahoo, load_material

    m = load_material("ER70S-6.v1")
    knots = m.tables.dgamma_dT
    assert len(knots) == 8
    assert knots[0][0] == m.T_liquidus and knots[-1][0] == max(m.T_liquidus + 800.0, 2600.0)
    for T, dg in knots:
        h = 1e-3
        fd = (gamma_sahoo(T + h, 28.0, 1793.0) - gamma_sahoo(T - h, 28.0, 1793.0)) / (2 * h)
        assert dg == pytest.approx(fd, rel=1e-5, abs=1e-9)
        assert dg == dgamma_dT_sahoo(T, 28.0)
    # sulphur-free melt: the isotherm reduces to the constant -A
    assert dgamma_dT_sahoo(2000.0, 0.0) == pytest.approx(-4.3e-4)
    assert m.dgamma_dT == dgamma_dT_sahoo(m.T_liquidus + 50.0, 28.0)


def test_torch_path_csv_resolves_against_the_package():
    from waam_solver_amd.job import parse_torch_path

    pts = parse_torch_path({"torch_path_csv": "jobs/paths/bead_calibrate.csv",
                            "_base_dir": "/nonexistent"})
    assert pts == [(0.015, 0.02, 0.0), (0.055, 0.02, 0.0)]
    pts = parse_torch_path({"torch_path": [{"x_mm": 5, "y_mm": 4}, {"x_mm": 9, "y_mm": 4}]})
    assert pts == [(0.005, 0.004, 0.0), (0.009, 0.004, 0.0)]


@pytest.fixture(scope="module")
def bead_twin():
    from waam_solver_amd.job import PKG_ROOT
    from waam_solver_amd.twin import WAAMTwin

    return WAAMTwin.from_job(PKG_ROOT / BEAD_JOB, device="cpu")


def test_bead_on_plate_job_settings(bead_twin):
    tw = bead_twin
    g = tw.grid
    assert (g.nx, g.ny, g.nz) == (200, 100, 62) and g.dx == pytest.approx(4e-4)
    assert tw.use_material_tables and tw.gpu_tables.struct.n_dgamma == 8
    assert tw.Q_w == pytest.approx(700.0) and tw.eta == pytest.approx(0.72)
    assert tw.T_amb == pytest.approx(293.0)
    assert tw.travel_speed_m_s == pytest.approx(7.5e-3)
    assert tw.wire_feed_m_s == pytest.approx(6.5 / 60.0)
    assert tw.h_conv == pytest.approx(35.0) and tw.enable_radiation and tw.enable_convection
    assert tw.enable_vof and tw.enable_bead_freeze and not tw.enable_lorentz
    src = tw.arc_source
    a_f, a_r, b, c = src.resolved_axes(tw)
    assert (a_f, a_r, b, c) == pytest.approx((2.2 / 0.4, 4.2 / 0.4, 1.5 / 0.4, 1.0 / 0.4))
    from waam_solver_amd.physics.weld_forces import arc_pressure_peak_pa

    # Lin & Eagar: mu0 I^2 / (4 pi^2 sigma^2) with sigma the 2 mm arc sigma
    assert arc_pressure_peak_pa(tw) == pytest.approx(
        4e-7 * math.pi * 70.0 ** 2 / (4 * math.pi ** 2 * 2e-3 ** 2))


def test_calibration_overlay_from_job(tmp_path):
    import yaml

    from waam_solver_amd.job import PKG_ROOT
    from waam_solver_amd.twin import WAAMTwin

    job = yaml.safe_load(open(PKG_ROOT / BEAD_JOB))
    job["calibration"] = "materials/calibration/ER70S-6.bead_on_plate.yaml"
    job["simulation"]["domain_mm"] = [12, 8, 6]
    job["plate"] = {"thickness_mm": 2.0}
    p = tmp_path / "job.yaml"
    p.write_text(yaml.safe_dump(job))
    tw = WAAMTwin.from_job(p, device="cpu")
    assert tw.eta == pytest.approx(0.65)
    assert tw.sigma_cells == pytest.approx(2e-3 / 4e-4 * 0.8)


def test_oracle_twin_like_a_job_twin(bead_twin):
    """The oracle twin built from a job twin derives the same constants and host quantities."""
    from oracle.oracle import pack_tables

    tw = bead_twin
    ot = oracle_like(tw)
    a, b = derived_constants(tw), derived_constants(ot)
    for k in a:
        assert a[k] == b[k], k
    assert ot.use_tables and ot.H_sol == tw.H_sol and ot.H_liq == tw.H_liq
    knots, counts = pack_tables(ot.mat)
    s = tw.gpu_tables.struct
    got = np.concatenate([np.array(getattr(s, f"{p}_{tv}"), np.float32)
                          for p in ("cp", "k", "mu", "dgamma") for tv in ("T", "V")])
    np.testing.assert_array_equal(got, knots)
    assert list(counts) == [s.n_cp, s.n_k, s.n_mu, s.n_dgamma]
    from waam_solver_amd.physics import deposition_balance as db
    from waam_solver_amd.physics.electrical_stickout import droplet_entry_temperature_K
    from waam_solver_amd.physics.weld_forces import (arc_pressure_peak_pa,
                                                     arc_pressure_sigma_cells,
                                                     droplet_impact_velocity_m_s)

    assert ot.arc_pressure_peak_pa() == arc_pressure_peak_pa(tw)
    assert ot.arc_pressure_sigma_m() / ot.g.dx == arc_pressure_sigma_cells(tw)
    assert ot.droplet_period_s() == db.droplet_period_s(tw)
    assert ot.droplet_mass_kg() == db.droplet_mass_kg(tw)
    assert ot.droplet_entry_temperature_K() == droplet_entry_temperature_K(tw)
    m = ot.droplet_mass_kg()
    assert ot.droplet_impact_velocity_m_s(m) == droplet_impact_velocity_m_s(tw, m)
    assert ot.resolve_plate_ij() == tw.resolve_plate_ij()
    assert ot.goldak == (0.6, 1.4, *tw.arc_source.resolved_axes(tw))
