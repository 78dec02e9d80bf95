"""Pins of the oracle's export-side restatements (CPU): vorticity against the curl of a rigid
rotation, the toe angle against its defining slope on a constructed column, and the force
diagnostics against closed forms of the forces it isolates."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.oracle import FLUID, GAS, SOLID, lib, ptr


def test_vorticity_of_a_rigid_rotation():
    """u = (-w y, w x, 0) has curl (0, 0, 2w); central differences are exact on a linear field,
    so interior cells hold 2w/dt and the clamped faces in x and y half of each term there; GAS
    cells are 0 (reference kernels.py:3380-3419)."""
    nx, ny, nz = 9, 8, 5
    w, dt = 0.01, 2e-5
    x, y, _ = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ux = (-w * y).astype(np.float32)
    uy = (w * x).astype(np.float32)
    uz = np.zeros_like(ux)
    flags = np.full((nx, ny, nz), FLUID, np.uint8)
    flags[:, :, -1] = GAS
    out = np.full((nx, ny, nz), -1.0, np.float32)
    lib().or_compute_vorticity_magnitude(ptr(ux), ptr(uy), ptr(uz), ptr(flags), ptr(out), dt,
                                         nx, ny, nz)
    scale = 1.0 / (dt + 1e-12)
    np.testing.assert_allclose(out[1:-1, 1:-1, :-1], 2 * w * scale, rtol=2e-6)
    # x faces: d(uy)/dx is a one-sided half difference, d(ux)/dy stays central
    np.testing.assert_allclose(out[0, 1:-1, :-1], 1.5 * w * scale, rtol=2e-6)
    np.testing.assert_allclose(out[1:-1, 0, :-1], 1.5 * w * scale, rtol=2e-6)
    np.testing.assert_allclose(out[0, 0, :-1], 1.0 * w * scale, rtol=2e-6)
    assert np.all(out[:, :, -1] == 0.0)


def _column_twin(phi_col, nz_solid=4):
    from oracle.oracle import Material, OracleTwin
    from tests.fixtures import ER70S6_PLACEHOLDER

    ot = OracleTwin(Material(**ER70S6_PLACEHOLDER), nx=9, ny=6, nz=len(phi_col), dx=2e-4,
                    max_tracers=8)
    ot.reset()
    ot.nz_solid = nz_solid
    ot.g.phi[ot.g.nx // 3, ot.g.ny // 2, :] = np.asarray(phi_col, np.float32)
    return ot


def test_toe_angle_follows_the_last_steep_step():
    # cells nz_solid-1 .. nz_solid+7 = 3..11; the last k > 3 with phi > 0.45 and a step is k=6
    col = [1, 1, 1, 1, 1.0, 0.9, 0.5, 0.2, 0.0, 0, 0, 0, 0, 0]
    ot = _column_twin(col)
    dphi = np.float32(0.5) - np.float32(0.9)
    want = math.degrees(math.atan(1.0 / abs(float(dphi))))
    assert ot.toe_angle_deg() == pytest.approx(min(89.0, max(10.0, want)), rel=1e-6)
    # a step of 0.02 gives atan(50) = 88.9 deg; a flat column keeps the contact angle
    col2 = [1, 1, 1, 1, 1.0, 0.52, 0.5, 0, 0, 0, 0, 0, 0, 0]
    assert _column_twin(col2).toe_angle_deg() == pytest.approx(
        math.degrees(math.atan(1.0 / float(abs(np.float32(0.5) - np.float32(0.52))))), rel=1e-6)
    ot3 = _column_twin([1.0] * 14)
    assert ot3.toe_angle_deg() == ot3.contact_angle_deg
    # a step just above the 1e-4 threshold gives atan(2000) = 89.97 deg, clamped to 89
    ot4 = _column_twin([1, 1, 1, 1, 1.0, 0.5005, 0.5, 0, 0, 0, 0, 0, 0, 0])
    assert ot4.toe_angle_deg() == 89.0


def test_force_diagnostics_isolate_each_force():
    """On a still, isothermal pool at T_liquidus with f_l = 1 and no interface, CSF and
    Marangoni vanish, buoyancy is zero (T = T_l) and the hydrostatic term alone is g_lu; the
    restored live forces come back unchanged."""
    from oracle.oracle import Material, OracleTwin
    from tests.fixtures import ER70S6_PLACEHOLDER

    ot = OracleTwin(Material(**ER70S6_PLACEHOLDER), nx=10, ny=8, nz=8, dx=2e-4, max_tracers=8,
                    enable_vof=True, enable_csf_tension=True, enable_hydrostatic_gravity=True)
    ot.reset()
    g = ot.g
    g.flags[...] = FLUID
    g.flags[:, :, 0] = SOLID
    g.phi[...] = 1.0
    g.f_l[...] = 1.0
    g.T[...] = np.float32(ot.mat.T_liquidus)
    rng = np.random.default_rng(1)
    live = [rng.standard_normal(g.Fx.shape).astype(np.float32) for _ in range(3)]
    g.Fx[...], g.Fy[...], g.Fz[...] = live
    ot.last_arc_ijk = (5.0, 4.0, 7.0)
    d = ot.force_diagnostics()
    assert d["f_csf_max"] == 0.0 and d["f_marangoni_max"] == 0.0
    assert d["f_buoyancy_max"] == 0.0
    assert d["f_gravity_max"] == round(float(np.float32(ot.g_lu)), 6)
    assert d["f_arc_max"] == 0.0  # no interface cell under the arc
    want = float(np.sqrt(live[0] ** 2 + live[1] ** 2 + live[2] ** 2).max())
    assert d["f_total_snap_max"] == round(want, 6)
    for a, b in zip((g.Fx, g.Fy, g.Fz), live):
        np.testing.assert_array_equal(a, b)
