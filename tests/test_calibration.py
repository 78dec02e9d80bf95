"""Calibration overlays on a twin (host-side; the twin is built on the CPU device and no kernel
runs)."""

from __future__ import annotations

import pytest

from waam_solver_amd.calibration import CalibrationProfile, apply_calibration, load_calibration


@pytest.fixture()
def twin():
    from waam_solver_amd.twin import WAAMTwin

    return WAAMTwin(material="LCS-bead-on-plate", nx=16, ny=8, nz=8, device="cpu")


def test_identity_profile_changes_nothing(twin):
    before = (twin.eta, twin.h_conv, twin.marangoni_scale, twin.sigma_cells)
    apply_calibration(twin, CalibrationProfile())
    assert (twin.eta, twin.h_conv, twin.marangoni_scale, twin.sigma_cells) == before
    assert load_calibration(None) == CalibrationProfile()
    assert load_calibration("") == CalibrationProfile()


def test_overlay_semantics(twin, tmp_path):
    h0, s0 = twin.h_conv, twin.sigma_cells
    f = tmp_path / "cal.yaml"
    f.write_text("arc_efficiency: 0.75\nheat_loss_factor: 2.0\narc_sigma_scale: 1.1\n")
    cal = load_calibration(str(f))
    assert cal == CalibrationProfile(0.75, 2.0, 1.0, 1.1)
    apply_calibration(twin, cal)
    assert twin.eta == 0.75            # replaced, not multiplied
    assert twin.h_conv == h0 * 2.0     # multiplied
    assert twin.sigma_cells == s0 * 1.1
    assert twin.marangoni_scale == 1.0
    apply_calibration(twin, cal)       # a second overlay multiplies again, as in the reference
    assert twin.h_conv == h0 * 4.0
