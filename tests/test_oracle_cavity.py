"""The oracle's lid-driven cavity step (CPU) against closed forms: plane Couette flow between a
fixed wall and the moving lid (linear profile, walls half a cell outside the fluid), the rest
state as a fixed point, the fluid's mass and the mirror symmetry of the cavity.

The collision is the reference's central-moment map, whose second-order reconstruction adds
e e : D without removing its trace (sum of the post-collision populations = rho + 1.5 tr D):
under shear it neither conserves mass exactly nor reproduces Couette flow to rounding. The
bounds below are that defect's size (measured: 5.6e-5 of u_w on the Couette profile at
tau = 1, 1e-4 relative mass over 1000 cavity steps); a dropped or mis-signed moving-wall term,
or a wall half a cell off, moves the profile by 5 % or more."""

from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import lib, ptr

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float32)


def _run(flags, f, omega, u_lid, steps):
    nx, ny, nz = flags.shape
    L = lib()
    a, b = f.copy(), f.copy()
    rho, ux, uy, uz = (np.zeros(flags.shape, np.float32) for _ in range(4))
    for _ in range(steps):
        L.or_cavity_step(ptr(a), ptr(b), ptr(rho), ptr(ux), ptr(uy), ptr(uz), ptr(flags), omega,
                         omega, u_lid, 0.0, 0.0, nx, ny, nz)
        a, b = b, a
    return a, rho, ux, uy, uz


def test_couette_profile_is_linear_between_the_walls():
    """Periodic in x and y, a wall layer at k = 0 and the lid at k = nz - 1: the steady flow is
    u_x(k) = u_w (k - 1/2) / H with H = nz - 2 fluid layers (half-way bounce-back puts each wall
    half a cell outside the fluid); nothing else moves."""
    from waam_solver_amd.cavity import FLUID, LID, WALL

    nx, ny, nz = 4, 3, 12
    flags = np.full((nx, ny, nz), FLUID, np.uint8)
    flags[:, :, 0] = WALL
    flags[:, :, -1] = LID
    f = np.repeat(W[:, None], nx * ny * nz, 1).reshape(19, nx, ny, nz).astype(np.float32)
    uw = 0.05
    _, rho, ux, uy, uz = _run(flags, f, 1.0, uw, 3000)
    H = nz - 2
    want = uw * (np.arange(1, nz - 1) - 0.5) / H
    got = ux[:, :, 1:-1]
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=0, atol=1e-4 * uw)
    assert np.abs(uy[:, :, 1:-1]).max() < 1e-6 and np.abs(uz[:, :, 1:-1]).max() < 1e-6
    np.testing.assert_allclose(rho[:, :, 1:-1], 1.0, atol=2e-5)


def test_rest_state_is_a_fixed_point():
    from waam_solver_amd.cavity import cavity_flags

    flags = cavity_flags(8, 8, 8)
    f = np.repeat(W[:, None], 512, 1).reshape(19, 8, 8, 8).astype(np.float32)
    a, rho, ux, uy, uz = _run(flags, f, 1.0 / 0.596, 0.0, 40)
    fluid = flags == 0
    assert np.all(ux[fluid] == 0) and np.all(uy[fluid] == 0) and np.all(uz[fluid] == 0)
    # the 19 weights do not sum to 1 exactly in fp32, so the populations settle within an ulp
    np.testing.assert_allclose(a[:, fluid], f[:, fluid], rtol=1e-6, atol=0)


def test_mass_is_conserved_and_the_flow_mirror_symmetric():
    """Bounce-back returns every population that leaves a fluid cell and the lid's momentum
    terms cancel in pairs (+-e_x across the lid), so the fluid's mass moves only by the
    collision's trace defect. Without density noise the flow is symmetric under y -> -y (u_y
    odd, u_x and u_z even) up to rounding (mirrored cells add the same values in another
    order)."""
    from waam_solver_amd.cavity import cavity_flags, initial_populations

    n = 12
    flags = cavity_flags(n + 2, n + 2, n + 2)
    fluid = flags == 0
    f, _ = initial_populations(n + 2, n + 2, n + 2, 1e-4, 3)
    m0 = f[:, fluid].astype(np.float64).sum()
    a, *_ = _run(flags, f, 1.0 / 0.62, 0.05, 300)
    m1 = a[:, fluid].astype(np.float64).sum()
    assert abs(m1 - m0) / m0 < 2e-4
    f0, _ = initial_populations(n + 2, n + 2, n + 2, 0.0, 3)
    _, rho, ux, uy, uz = _run(flags, f0, 1.0 / 0.62, 0.05, 300)
    assert np.abs(ux).max() > 0.005
    np.testing.assert_allclose(uy, -uy[:, ::-1, :], atol=1e-6)
    np.testing.assert_allclose(ux, ux[:, ::-1, :], atol=1e-6)
    np.testing.assert_allclose(uz, uz[:, ::-1, :], atol=1e-6)
