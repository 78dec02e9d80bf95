"""The oracle reproduces the reference's own validation gates and measured values.

Each test rebuilds one reference validation case on the CPU oracle and checks it against the
gate the reference enforces and, where the reference records one, the value it reports as
measured (reference README.md:666-676 and the test-file comments cited per test). Together
these pin the restated kernels end to end: a dropped term, a wrong sign or a wrong buffer
choreography moves these numbers.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.oracle import GAS, IFACE, FLUID, SOLID, Material, OracleTwin, ptr
from tests.fixtures import (ER70S6_PLACEHOLDER, f32, gaussian_pulse, poiseuille_channel,
                            soft_sphere_phi, stefan_interface_height, stefan_lambda)


def _twin(**kw) -> OracleTwin:
    return OracleTwin(Material(**ER70S6_PLACEHOLDER), **kw)


def test_poiseuille_gate_reproduces_dead_stream_value():
    """validation/test_lbm_poiseuille.py:16-87 — gate 8 %, measured 6.5 %.

    With the reference's buffer choreography the fluid cells relax locally. At omega = 1 the
    Guo-forced BGK step maps the post-collision momentum j to j + F, and the stored velocity
    is (j + F/2)/rho, so after n steps every fluid cell carries u0 + (n - 1/2) F: a uniform
    shift of 99.5*fx after 100 steps, which predicts a normalised profile error of 6.511 %
    (the reference reports "measured 6.5%").
    """
    nx, ny, nz = 64, 32, 6
    tau = 1.0
    nu = (tau - 0.5) / 3.0
    fx = 1e-4
    t = _twin(nx=nx, ny=ny, nz=nz, dx=3e-4, C_darcy=0.0, max_tracers=10)
    g, L = t.g, t.lib
    fix = poiseuille_channel(nx, ny, nz, fx, nu)
    g.f_a[...] = fix["f"]
    g.flags[...] = fix["flags"]
    g.ux[...] = fix["ux"]
    g.rho[...] = fix["rho"]
    g.f_l[...] = 1.0
    g.phi[...] = 1.0
    L.or_stream(ptr(g.f_a), ptr(g.f_b), ptr(g.flags), nx, ny, nz)
    fluid = g.flags == FLUID
    for _ in range(100):
        g.Fx[fluid] = f32(fx)
        g.Fy[fluid] = 0.0
        g.Fz[fluid] = 0.0
        L.or_collide_srt(ptr(g.f_src), ptr(g.f_dst), ptr(g.rho), ptr(g.ux), ptr(g.uy),
                         ptr(g.uz), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.f_l), ptr(g.flags),
                         tau, 1.0 / tau, 1.0, 0.0, nx, ny, nz)
        L.or_stream(ptr(g.f_dst), ptr(g.f_src), ptr(g.flags), nx, ny, nz)
        g.swap_buffers()
    mi, mk = nx // 2, nz // 2
    u_meas = np.array([g.ux[mi, j, mk] for j in range(1, ny - 1) if g.flags[mi, j, mk] == FLUID],
                      dtype=np.float64)
    y = np.arange(1, ny - 1, dtype=np.float64) - 1.0
    u_an = (fx / (2.0 * nu)) * y * ((ny - 2) - 1.0 - y)
    l2 = float(np.sqrt(np.mean((u_meas / u_meas.max() - u_an / u_an.max()) ** 2)) * 100.0)
    # closed form of the dead-stream prediction: profile + uniform 99.5*fx shift
    u_pred = u_an + 99.5 * fx
    l2_pred = float(np.sqrt(np.mean((u_pred / u_pred.max() - u_an / u_an.max()) ** 2)) * 100.0)
    assert l2 < 8.0
    assert abs(l2_pred - 6.511) < 0.001
    assert abs(l2 - l2_pred) < 0.002, (l2, l2_pred)


def test_thermal_diffusion_gate():
    """validation/test_thermal_diffusion.py:17-78 — gate 1 %, measured 0.11 %."""
    n, dx, T_bg, A, sig = 48, 3e-4, 300.0, 500.0, 1.2e-3
    t = _twin(nx=n, ny=n, nz=n, dx=dx, use_srt=True, C_darcy=0.0, max_tracers=100)
    t.reset(T_ambient=T_bg, test_fluid_domain=True)
    g, L = t.g, t.lib
    r2, inv2s2 = gaussian_pulse(n, dx, T_bg, A, sig, t.cp_rho)
    T0 = (f32(T_bg) + f32(A) * np.exp(-r2 * inv2s2)).astype(np.float32)
    g.T[...] = T0
    g.H[...] = f32(t.cp_rho) * T0
    g.f_l[...] = 1.0
    H_in = np.empty_like(g.H)
    for _ in range(400):
        H_in[...] = g.H
        L.or_advect_diffuse_temperature_jacobi(ptr(H_in), ptr(g.H), ptr(g.T), ptr(g.ux),
                                               ptr(g.uy), ptr(g.uz), ptr(g.flags), t.alpha_lu,
                                               1.0, t.cp_rho, n, n, n)
        g.T[...] = g.H / f32(t.cp_rho)
    tt = 400 * g.dt
    x = (np.arange(n) - n // 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    R2 = X ** 2 + Y ** 2 + Z ** 2
    s2 = sig ** 2 + 2.0 * t.mat.alpha * tt
    T_an = T_bg + A * (sig ** 2 / s2) ** 1.5 * np.exp(-R2 / (2.0 * s2))
    sl = np.s_[4:-4, 4:-4, 4:-4]
    l2 = float(np.sqrt(np.sum((g.T[sl] - T_an[sl]) ** 2) / np.sum((T_an[sl] - T_bg) ** 2)) * 100)
    assert l2 < 1.0
    assert abs(l2 - 0.11) < 0.01, l2


def test_thermal_serial_equals_jacobi_without_flow():
    """With u = 0 the thermal kernel reads no neighbouring H, so both orders agree bitwise."""
    n = 12
    t = _twin(nx=n, ny=n, nz=n, dx=3e-4, max_tracers=10)
    t.reset(test_fluid_domain=True)
    g, L = t.g, t.lib
    rng = np.random.default_rng(3)
    g.T[...] = (300 + 500 * rng.random(g.T.shape)).astype(np.float32)
    g.H[...] = f32(t.cp_rho) * g.T
    g.flags[rng.random(g.T.shape) < 0.2] = GAS
    H0 = g.H.copy()
    a = H0.copy()
    L.or_advect_diffuse_temperature_serial(ptr(a), ptr(g.T), ptr(g.ux), ptr(g.uy), ptr(g.uz),
                                           ptr(g.flags), t.alpha_lu, 1.0, t.cp_rho, n, n, n)
    b = np.empty_like(H0)
    L.or_advect_diffuse_temperature_jacobi(ptr(H0), ptr(b), ptr(g.T), ptr(g.ux), ptr(g.uy),
                                           ptr(g.uz), ptr(g.flags), t.alpha_lu, 1.0, t.cp_rho,
                                           n, n, n)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_stefan_gate():
    """validation/test_stefan_solidification.py:82-163 — gate 8 %, measured 6.0 %."""
    nx, ny, nz, nzs, dx = 4, 4, 80, 8, 1e-4
    t = _twin(nx=nx, ny=ny, nz=nz, dx=dx, C_darcy=0.0, max_tracers=10)
    g, L, m = t.g, t.lib, t.mat
    T_amb, T_init = 300.0, m.T_liquidus + 50.0
    cp, Lr = f32(t.cp_rho), f32(t.L_rho)
    H_liq = cp * f32(m.T_liquidus) + Lr
    g.flags[:, :, :nzs] = SOLID
    g.flags[:, :, nzs:] = FLUID
    g.phi[...] = 1.0
    g.f_l[:, :, :nzs] = 0.0
    g.f_l[:, :, nzs:] = 1.0
    g.T[:, :, :nzs] = T_amb
    g.T[:, :, nzs:] = T_init
    g.H[:, :, :nzs] = cp * f32(T_amb)
    g.H[:, :, nzs:] = H_liq + cp * (f32(T_init) - f32(m.T_liquidus))
    ste = m.cp * (T_init - T_amb) / m.L_fusion
    slope_an = 2.0 * stefan_lambda(ste) * math.sqrt(m.alpha)
    samples, st, ss = [200, 300, 400, 500, 600, 700], [], []
    H_in = np.empty_like(g.H)
    for step in range(1, 801):
        sub = (np.arange(nz) < nzs)[None, None, :] & (g.flags == SOLID)
        g.H[sub] = cp * f32(T_amb)
        g.T[sub] = f32(T_amb)
        g.f_l[sub] = 0.0
        H_in[...] = g.H
        L.or_advect_diffuse_temperature_jacobi(ptr(H_in), ptr(g.H), ptr(g.T), ptr(g.ux),
                                               ptr(g.uy), ptr(g.uz), ptr(g.flags), t.alpha_lu,
                                               1.0, t.cp_rho, nx, ny, nz)
        L.or_update_phase(ptr(g.H), ptr(g.T), ptr(g.f_l), t.cp_rho, t.L_rho, m.T_solidus,
                          m.T_liquidus, nx, ny, nz)
        if step in samples:
            st.append(math.sqrt(step * g.dt))
            ss.append(stefan_interface_height(g.f_l, g.T, g.flags, nzs, dx, m.T_solidus,
                                              m.T_liquidus))
    assert all(ss[i] <= ss[i + 1] for i in range(len(ss) - 1))
    slope = float(np.polyfit(st, ss, 1)[0])
    err = abs(slope - slope_an) / slope_an * 100.0
    assert err < 8.0
    assert abs(err - 6.0) < 0.1, err


def test_vof_mass_gate():
    """validation/test_vof_mass.py:14-41 — gate 0.5 %, measured 0.00 %."""
    t = _twin(nx=32, ny=32, nz=16, dx=3e-4, C_darcy=0.0, max_tracers=10, enable_vof=True)
    t.reset(test_fluid_domain=True)
    g, L = t.g, t.lib
    phi0 = float(g.phi.sum())
    g.ux[g.flags == FLUID] = f32(0.02)
    for _ in range(80):
        L.or_advect_phi(ptr(g.phi_tmp), ptr(g.phi), ptr(g.ux), ptr(g.uy), ptr(g.uz),
                        ptr(g.flags), 32, 32, 16)
        g.phi[...] = g.phi_tmp
        L.or_reinitialize_phi(ptr(g.phi), ptr(g.flags), 32, 32, 16)
    drift = abs(float(g.phi.sum()) - phi0) / phi0 * 100.0
    assert drift < 0.5 and drift == 0.0


def test_vof_mass_conserved_for_a_blob():
    """Flux-form donor-cell advection conserves sum(phi) to round-off for any phi field."""
    n = 16
    t = _twin(nx=n, ny=n, nz=n, dx=3e-4, max_tracers=10, enable_vof=True)
    t.reset(test_fluid_domain=True)
    g, L = t.g, t.lib
    rng = np.random.default_rng(5)
    g.phi[...] = soft_sphere_phi(n, 4.0)
    g.ux[...] = f32(0.03)
    g.uy[...] = (0.02 * rng.standard_normal(g.uy.shape)).astype(np.float32)
    g.uy[:, 0, :] = 0.0
    g.uy[:, -1, :] = 0.0
    s0 = float(g.phi.astype(np.float64).sum())
    for _ in range(20):
        L.or_advect_phi(ptr(g.phi_tmp), ptr(g.phi), ptr(g.ux), ptr(g.uy), ptr(g.uz),
                        ptr(g.flags), n, n, n)
        g.phi[...] = g.phi_tmp
    s1 = float(g.phi.astype(np.float64).sum())
    assert abs(s1 - s0) / s0 < 1e-5


def test_mass_conservation_gate():
    """validation/test_mass_conservation.py:16-48 — rho drift < 2 % over 500 SRT steps."""
    t = _twin(nx=32, ny=32, nz=16, dx=3e-4, C_darcy=0.0, max_tracers=50)
    t.reset(test_fluid_domain=True)
    g, L = t.g, t.lib
    rho0 = float(g.rho.mean())
    for _ in range(500):
        g.Fx[...] = 0
        g.Fy[...] = 0
        g.Fz[...] = 0
        L.or_collide_srt(ptr(g.f_src), ptr(g.f_dst), ptr(g.rho), ptr(g.ux), ptr(g.uy),
                         ptr(g.uz), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.f_l), ptr(g.flags),
                         g.tau, t.omega, 1.0, 0.0, 32, 32, 16)
        L.or_stream(ptr(g.f_dst), ptr(g.f_src), ptr(g.flags), 32, 32, 16)
        g.swap_buffers()
    drift = abs(float(g.rho.mean()) - rho0) / rho0 * 100
    assert drift < 2.0


def test_laplace_curvature_gate():
    """validation/test_laplace.py:31-103 — Brackbill kappa within 15 % of 2/R at R = 8."""
    n, R = 40, 8.0
    t = _twin(nx=n, ny=n, nz=n, dx=2.5e-4, enable_csf_tension=True, enable_vof=True,
              C_darcy=0.0, max_tracers=10)
    t.reset(test_fluid_domain=True)
    g, L = t.g, t.lib
    phi = soft_sphere_phi(n, R)
    flags = np.full((n, n, n), FLUID, np.uint8)
    flags[phi < 0.05] = GAS
    flags[(phi >= 0.05) & (phi <= 0.95)] = IFACE
    g.phi[...] = phi
    g.flags[...] = flags
    kappa = np.zeros_like(phi)
    L.or_compute_curvature_field(ptr(g.phi), ptr(g.flags), ptr(kappa), n, n, n)
    iface = (phi > 0.25) & (phi < 0.75) & (flags != GAS)
    iface[:2], iface[-2:], iface[:, :2], iface[:, -2:] = False, False, False, False
    iface[:, :, :2], iface[:, :, -2:] = False, False
    km = abs(float(kappa[iface].mean()))
    assert abs(km - 2.0 / R) / (2.0 / R) <= 0.15
    g.Fx[...] = 0
    g.Fy[...] = 0
    g.Fz[...] = 0
    L.or_compute_csf_tension(ptr(g.phi), ptr(g.flags), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz),
                             t.gamma_lu, 0, 0.0, 1.0, n, n, n)
    assert float(np.sqrt(g.Fx ** 2 + g.Fy ** 2 + g.Fz ** 2).max()) > 1e-12


@pytest.mark.parametrize("source", ["goldak", "gaussian2d"])
def test_arc_energy_gate(source):
    """validation/test_goldak_energy.py:16-69 — injected enthalpy = eta*Q*dt within 1 %."""
    t = _twin(nx=32, ny=24, nz=20, dx=3e-4, heat_source=source, arc_efficiency=0.8,
              arc_power_W=3600.0, max_tracers=10)
    t.reset(test_fluid_domain=True)
    t.goldak = (0.6, 1.4, 4.0, 8.0, 4.0, 3.0)
    g = t.g
    H0 = float(g.H.astype(np.float64).sum()) * g.dx ** 3
    ai, aj, ak = (g.nx - 1) * 0.5, (g.ny - 1) * 0.5, float(t.nz_solid + 2)
    t._inject_arc(ai, aj, ak)
    dH = float(g.H.astype(np.float64).sum()) * g.dx ** 3 - H0
    exp_ = t.eta * t.Q_w * g.dt
    assert abs(dH - exp_) / exp_ <= 0.01
    if source == "goldak":
        i0 = int(round(ai))
        assert float(g.H[:i0].sum()) > float(g.H[i0 + 1:].sum())


def test_marangoni_direction_gate():
    """validation/test_marangoni_cell.py:16-63 — dgamma/dT < 0 pulls from hot to cold."""
    t = _twin(nx=40, ny=16, nz=20, dx=2.5e-4, max_tracers=10)
    t.reset()
    g, L = t.g, t.lib
    nzs, j = t.nz_solid, g.ny // 2
    for i in range(g.nx):
        for k in range(g.nz):
            if k < nzs - 1:
                g.phi[i, j, k], g.f_l[i, j, k], g.flags[i, j, k] = 1.0, 0.0, SOLID
            elif k == nzs - 1:
                g.phi[i, j, k], g.f_l[i, j, k], g.flags[i, j, k] = 0.5, 1.0, IFACE
                g.T[i, j, k] = 300.0 + 200.0 * (i / (g.nx - 1))
            else:
                g.phi[i, j, k], g.f_l[i, j, k], g.flags[i, j, k] = 0.0, 0.0, GAS
    L.or_compute_marangoni_force(ptr(g.T), ptr(g.phi), ptr(g.f_l), ptr(g.Fx), ptr(g.Fy),
                                 ptr(g.Fz), ptr(g.flags), t.dgamma_dT_lu, g.nx, g.ny, g.nz)
    Fx = g.Fx[:, j, nzs - 1]
    dT = np.gradient(g.T[:, j, nzs - 1])
    assert Fx[dT > 0].mean() < 0.0
