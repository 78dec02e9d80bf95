"""The differentiation form on the GPU: the thermal rollout (frozen flags, zero velocity) against
the oracle, its adjoint gradients against central finite differences of the oracle's binary64
build (h = 1e-4), checkpoint invariance, and a calibration fit by gradient descent.

The 128x64x48 case runs 400 steps with checkpoints every 50 (the configuration the
differentiation form is specified on): bead-on-plate plate and seeded pool, Gaussian2D arc
moving +x at 8 mm/s, convective and radiative losses, enthalpy cap. Loss: mean squared error of
the final temperature (/100 K) and liquid fraction fields against a run at eta 0.75,
heat-loss factor 1.0, sigma scale 1.1."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THETA_TARGET = (0.75, 1.0, 1.1)


def _case(nx, ny, nz, hip_device):
    from tests.step_cases import build_pair, case_kwargs

    tw, _ = build_pair("bead_mrt", layout=0)
    if (tw.grid.nx, tw.grid.ny, tw.grid.nz) != (nx, ny, nz):
        from tests.step_cases import configure, fixture_arrays
        from tests.fixtures import LCS_BEAD_ON_PLATE
        from waam_solver_amd.materials import material_from_dict
        from waam_solver_amd.twin import WAAMTwin

        kw = dict(case_kwargs("bead_mrt"), nx=nx, ny=ny, nz=nz)
        mat_d = dict(LCS_BEAD_ON_PLATE)
        m = material_from_dict({"name": mat_d["name"], "status": "synthetic",
                                "constants": {k: v for k, v in mat_d.items() if k != "name"}})
        tw = WAAMTwin(material=m, field_layout=0, **kw)
        configure(tw, "bead_mrt")
        tw.reset()
        g = tw.grid
        T, fl, phi, flags, H = (g.T.to_numpy(), g.f_l.to_numpy(), g.phi.to_numpy(),
                                g.flags.to_numpy().astype(np.uint8), g.H.to_numpy())
        fixture_arrays(T, fl, phi, flags, H, tw.nz_solid, mat_d, tw.cp_rho, tw.L_rho)
        for n, a in (("T", T), ("f_l", fl), ("phi", phi), ("flags", flags), ("H", H)):
            getattr(g, n).from_numpy(a)
    g = tw.grid
    x0 = (g.nx // 3) * g.dx
    y0 = (g.ny // 2) * g.dx
    path = lambda n: (x0 + 8e-3 * n * g.dt, y0)  # noqa: E731
    return tw, path


def _oracle_rollout(tw, rollout, theta, f64=True):
    """The same thermal chain on the oracle (binary64 build by default), from the rollout's
    starting state; returns final (H, T, f_l) in (nx, ny, nz) order."""
    from oracle.oracle import lib, ptr

    g = tw.grid
    dt_np = np.float64 if f64 else np.float32
    L = lib(parallel=True, f64=f64)
    d = (g.nx, g.ny, g.nz)
    H = rollout.H0.to_numpy().astype(dt_np)
    T = rollout.T0.to_numpy().astype(dt_np)
    fl = rollout.fl0.to_numpy().astype(dt_np)
    flags = g.flags.to_numpy().astype(np.uint8)
    zu = np.zeros(d, dt_np)
    eta, s_h, s_sig = theta
    m = tw.mat
    for n in range(rollout.n_steps):
        x, y = rollout.path(n)
        L.or_inject_arc_heat(ptr(H), ptr(flags), ptr(fl), None, x / g.dx, y / g.dx,
                             rollout.arc_k, tw.Q_w, rollout.sigma0 * s_sig, g.dt, g.dx ** 3, eta,
                             tw.arc_penetration_m / g.dx, 1 if tw.arc_surface_weighting else 0,
                             *d)
        L.or_clamp_enthalpy_ceiling_scalar(ptr(H), ptr(flags), tw.cp_rho, m.T_solidus,
                                           m.T_liquidus, tw.T_vapor_cap_K, tw.L_rho, *d)
        H_in = H.copy()
        L.or_advect_diffuse_temperature_jacobi(ptr(H_in), ptr(H), ptr(T), ptr(zu), ptr(zu),
                                               ptr(zu), ptr(flags), tw.alpha_lu, 1.0, tw.cp_rho,
                                               *d)
        L.or_apply_thermal_boundary_losses(ptr(H), ptr(T), ptr(flags), tw.T_amb,
                                           rollout.h_conv0 * s_h, rollout.eps0 * s_h,
                                           1 if tw.enable_convection else 0,
                                           1 if tw.enable_radiation else 0, tw.cp_rho, g.dt,
                                           g.dx, tw.sigma_sb, *d)
        L.or_clamp_enthalpy_floor_scalar(ptr(H), tw.cp_rho, ptr(flags), tw.T_amb, *d)
        L.or_update_phase(ptr(H), ptr(T), ptr(fl), tw.cp_rho, tw.L_rho, m.T_solidus,
                          m.T_liquidus, *d)
        L.or_clamp_enthalpy_ceiling_scalar(ptr(H), ptr(flags), tw.cp_rho, m.T_solidus,
                                           m.T_liquidus, tw.T_vapor_cap_K, tw.L_rho, *d)
        L.or_update_phase(ptr(H), ptr(T), ptr(fl), tw.cp_rho, tw.L_rho, m.T_solidus,
                          m.T_liquidus, *d)
    return H, T, fl


def _np_loss(T, fl, T_t, fl_t):
    return float((((T - T_t) / 100.0) ** 2).mean() + ((fl - fl_t) ** 2).mean())


def test_rollout_matches_the_oracle(hip_device):
    from waam_solver_amd.adjoint import ThermalRollout

    print("[adjoint] forward parity", flush=True)
    tw, path = _case(64, 40, 32, hip_device)
    ro = ThermalRollout(tw, 60, path, checkpoint_every=20)
    theta = (0.8, 1.2, 0.9)
    T, fl = ro.forward(theta)
    H, To, flo = _oracle_rollout(tw, ro, theta, f64=False)
    from tests.gpu_common import bits_equal, mismatch_report

    shape = (tw.grid.nx, tw.grid.ny, tw.grid.nz)
    for name, a, b in (("H", ro.H_final.cpu().numpy().reshape(shape), H),
                       ("T", T.cpu().numpy().reshape(shape), To),
                       ("f_l", fl.cpu().numpy().reshape(shape), flo)):
        assert bits_equal(a, b), mismatch_report(name, a, b)


def _grad_and_fd(tw, path, steps, K, theta0, hr, hip_device):
    import torch

    from waam_solver_amd.adjoint import ThermalRollout, ThermalRolloutFn, field_loss

    ro = ThermalRollout(tw, steps, path, checkpoint_every=K)
    T_t, fl_t = ro.forward(THETA_TARGET)
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    T, fl = ThermalRolloutFn.apply(th, ro)
    field_loss(T, fl, T_t, fl_t).backward()
    grad = th.grad.numpy()
    shape = (tw.grid.nx, tw.grid.ny, tw.grid.nz)
    Tt64 = T_t.cpu().numpy().reshape(shape).astype(np.float64)
    flt64 = fl_t.cpu().numpy().reshape(shape).astype(np.float64)
    fd = np.zeros(3)
    for p in range(3):
        h = hr * theta0[p]
        up, dn = list(theta0), list(theta0)
        up[p] += h
        dn[p] -= h
        _, Tu, flu = _oracle_rollout(tw, ro, up)
        _, Td, fld = _oracle_rollout(tw, ro, dn)
        fd[p] = (_np_loss(Tu, flu, Tt64, flt64) - _np_loss(Td, fld, Tt64, flt64)) / (2 * h)
    return grad, fd


def test_adjoint_gradient_against_binary64_finite_differences(hip_device):
    """128x64x48. Over 50 steps the loss is smooth in theta and the adjoint meets the binary64
    central differences (h = 1e-4 of each parameter) to 1e-4 (measured 1e-5 .. 4e-5). Over the
    400 steps (checkpoints every 50) cells cross f_l = 0.55, where the arc's surface weight
    steps, so the loss carries small jumps and the h = 1e-4 differences straddle some (they move
    by 0.25 % between h = 1e-4 and 1e-3, while h = 1e-3 and 1e-2 agree to 1e-4): there the
    adjoint (a.e. derivative) is checked against h = 1e-3, to 1e-3 for eta and the sigma scale
    and 2e-2 for the heat-loss factor, whose gradient is two orders smaller."""
    print("[adjoint] gradient vs FD, 128x64x48, 50 steps", flush=True)
    tw, path = _case(128, 64, 48, hip_device)
    theta0 = (0.8, 1.3, 1.0)
    grad, fd = _grad_and_fd(tw, path, 50, 50, theta0, 1e-4, hip_device)
    rel = np.abs(grad - fd) / np.abs(fd)
    print(f"[adjoint] 50 steps: grad {grad} fd {fd} rel {rel}", flush=True)
    assert np.all(rel < 1e-4), (grad, fd, rel)
    print("[adjoint] gradient vs FD, 128x64x48, 400 steps", flush=True)
    grad, fd = _grad_and_fd(tw, path, 400, 50, theta0, 1e-3, hip_device)
    rel = np.abs(grad - fd) / np.abs(fd)
    print(f"[adjoint] 400 steps: grad {grad} fd {fd} rel {rel}", flush=True)
    assert rel[0] < 1e-3 and rel[2] < 1e-3 and rel[1] < 2e-2, (grad, fd, rel)


def test_checkpoint_interval_does_not_change_the_gradient(hip_device):
    import torch

    from waam_solver_amd.adjoint import ThermalRollout, ThermalRolloutFn, field_loss

    tw, path = _case(64, 40, 32, hip_device)
    grads = []
    for K in (7, 100):
        ro = ThermalRollout(tw, 100, path, checkpoint_every=K)
        T_t, fl_t = ro.forward(THETA_TARGET)
        th = torch.tensor((0.8, 1.0, 1.0), dtype=torch.float64, requires_grad=True)
        T, fl = ThermalRolloutFn.apply(th, ro)
        field_loss(T, fl, T_t, fl_t).backward()
        grads.append(th.grad.numpy().copy())
    np.testing.assert_array_equal(grads[0], grads[1])


def test_fit_calibration_recovers_the_target(hip_device):
    from waam_solver_amd.adjoint import ThermalRollout, fit_calibration

    print("[adjoint] calibration fit", flush=True)
    tw, path = _case(64, 40, 32, hip_device)
    ro = ThermalRollout(tw, 100, path, checkpoint_every=25)
    T_t, fl_t = ro.forward(THETA_TARGET)
    hist = fit_calibration(ro, T_t, fl_t, (0.85, 1.0, 1.0), iters=30, lr=0.01)
    th, ls = hist["theta"], hist["loss"]
    print(f"[adjoint] fit: loss {ls[0]:.3e} -> {ls[-1]:.3e}; theta {th[0]} -> {th[-1]}", flush=True)
    assert ls[-1] < 0.1 * ls[0]
    assert abs(th[-1][0] - 0.75) < 0.02 and abs(th[-1][2] - 1.1) < 0.05
