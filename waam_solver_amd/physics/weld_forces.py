"""Arc pressure and droplet impact helpers (reference physics/weld_forces.py:22-193)."""

from __future__ import annotations

import math

from .. import kernels
from .deposition_balance import droplet_mass_kg, infer_transfer_mode

MU0 = 4.0e-7 * math.pi


def lin_eagar_peak_pa(current_A: float, sigma_p_m: float) -> float:
    """Lin & Eagar (1986) peak arc pressure mu0 I^2 / (4 pi^2 sigma^2) [Pa]."""
    sig = max(float(sigma_p_m), 1.0e-6)
    i = max(float(current_A), 0.0)
    return MU0 * i * i / (4.0 * math.pi * math.pi * sig * sig)


def arc_pressure_sigma_m(twin) -> float:
    sig = getattr(twin, "pressure_sigma_m", None)
    if sig is not None and sig > 0.0:
        return float(sig)
    return float(twin.arc_sigma_m)


def arc_pressure_sigma_cells(twin) -> float:
    return arc_pressure_sigma_m(twin) / twin.grid.dx


def arc_pressure_peak_pa(twin) -> float:
    model = str(getattr(twin, "arc_pressure_model", "constant")).strip().lower()
    if model in ("lin_eagar", "lineagar", "lin-eagar"):
        return lin_eagar_peak_pa(twin.welding_current_A, arc_pressure_sigma_m(twin))
    return float(twin.arc_pressure)


def droplet_impact_velocity_m_s(twin, m_drop_kg: float | None = None) -> float:
    drop_len = 3e-3
    if twin.droplet_freq > 0 and twin.wire_feed_m_s > 0:
        drop_len = twin.wire_feed_m_s / twin.droplet_freq
    if m_drop_kg is None:
        m_drop_kg = droplet_mass_kg(twin)
    v_feed = max(twin.wire_feed_m_s, 0.0)
    v_grav = math.sqrt(max(0.0, 2.0 * 9.81 * drop_len))
    dia_ref = max(twin.wire_diameter_m, 1e-6)
    pinch = 0.35 * math.sqrt(max(twin.welding_current_A, 0.0) / 200.0) * math.sqrt(1.2e-3 / dia_ref)
    scale = {"globular": 0.78, "spray": 1.18, "pulsed": 1.05}.get(infer_transfer_mode(twin), 1.0)
    mass_term = 1.0
    if m_drop_kg > 0:
        mass_term = max(0.8, min(1.25, math.sqrt(droplet_mass_kg(twin) / m_drop_kg)))
    return min(5.0, scale * max(v_feed, v_grav + pinch) * mass_term)


def droplet_velocity_lu(twin, m_drop_kg: float | None = None) -> tuple[float, float, float]:
    g = twin.grid
    v = droplet_impact_velocity_m_s(twin, m_drop_kg)
    lead = math.radians(float(getattr(twin, "impact_lead_angle_deg", 0.0)))
    dxd, dyd, _ = getattr(twin, "_torch_dir_xyz", (1.0, 0.0, 0.0))
    dxy = math.sqrt(dxd * dxd + dyd * dyd)
    if dxy < 1e-9:
        dxd, dyd, dxy = 1.0, 0.0, 1.0
    dxd /= dxy
    dyd /= dxy
    v_xy = max(twin.travel_speed_m_s, abs(v) * math.tan(lead))
    v_xy = min(v * 0.65, v_xy)
    v_z = -math.sqrt(max(v * v - v_xy * v_xy, v * v * 0.20))
    return (v_xy * dxd * g.dt / g.dx, v_xy * dyd * g.dt / g.dx, v_z * g.dt / g.dx)


def droplet_impact_pressure_pa(twin, m_drop_kg: float | None = None,
                               drop_r_m: float | None = None) -> float:
    m = droplet_mass_kg(twin) if m_drop_kg is None else m_drop_kg
    if m <= 0:
        return 0.0
    v = droplet_impact_velocity_m_s(twin, m)
    r = drop_r_m if drop_r_m is not None else max(twin.grid.dx * 1.5, twin.wire_diameter_m * 0.5)
    return (0.5 * m * v * v) / max(math.pi * r * r, 1e-12)


def apply_droplet_impact(twin, g, arc_i, arc_j, arc_k, drop_r, drop_mass_kg=None) -> None:
    if not twin.enable_deposition_momentum:
        return
    vx, vy, vz = droplet_velocity_lu(twin, drop_mass_kg)
    kernels.feed_wire_momentum_impact(g.f_src, g.rho, g.ux, g.uy, g.uz, g.flags, g.f_l, arc_i,
                                      arc_j, arc_k, drop_r, vx, vy, vz, g.FLAG_GAS, g.nx, g.ny,
                                      g.nz)
    if twin.enable_droplet_impact_pressure:
        kernels.apply_droplet_impact_pressure(
            g.Fz, g.flags, g.phi, g.f_l, arc_i, arc_j, arc_k, drop_r,
            droplet_impact_pressure_pa(twin, drop_mass_kg, drop_r * g.dx), g.dt, g.dx,
            twin.mat.rho, g.FLAG_SOLID, g.FLAG_GAS)


def gas_shear_tau_pa(twin) -> float:
    """Shielding-gas shear 0.5 rho_Ar v^2 c, capped at 5 kPa (reference
    physics/weld_forces.py:120-125)."""
    rho_gas = 1.6  # argon, kg/m^3
    v = twin.gas_jet_velocity_m_s
    return min(0.5 * rho_gas * v * v * twin.gas_shear_coeff, 5000.0)


def apply_recoil(twin, g, arc_i: float, arc_j: float, arc_k: float) -> None:
    """Vapour recoil, Clausius-Clapeyron or constant-pressure form (reference
    physics/weld_forces.py:128-148)."""
    from .. import kernels

    if not twin.enable_recoil:
        return
    if twin.use_recoil_clausius_clapeyron:
        kernels.apply_vapor_recoil_clausius_clapeyron(
            g.Fz, g.T, g.phi, g.flags, arc_i, arc_j, arc_k, twin.sigma_cells, twin.P_vapor_ref_Pa,
            twin.T_boiling_K, twin.L_vapor_J_kg, twin.R_spec_vapor_J_kgK,
            float(getattr(twin, "recoil_accommodation", 0.54)), g.dt, g.dx, twin.mat.rho,
            g.FLAG_SOLID, g.FLAG_GAS)
    else:
        kernels.apply_vapor_recoil(g.Fz, g.T, g.phi, g.flags, arc_i, arc_j, arc_k,
                                   twin.sigma_cells, twin.recoil_pressure_pa, twin.mat.T_liquidus,
                                   g.dt, g.dx, twin.mat.rho, g.FLAG_SOLID, g.FLAG_GAS)


def apply_gas_shear(twin, g, arc_i: float, arc_j: float, arc_k: float) -> None:
    """Reference physics/weld_forces.py:151-161."""
    from .. import kernels

    if not twin.enable_gas_shear:
        return
    kernels.apply_gas_shear_stress(g.Fx, g.Fy, g.phi, g.flags, arc_i, arc_j, arc_k,
                                   twin.sigma_cells, gas_shear_tau_pa(twin), g.dt, g.dx,
                                   twin.mat.rho, g.FLAG_SOLID, g.FLAG_GAS)


def solve_lorentz(twin, g, arc_i: float, arc_j: float, arc_k: float) -> None:
    """J x B force of the welding current (reference physics/weld_forces.py:196-281).

    Conductivity from the phase state, a Gaussian current source normalised to the welding
    current, the potential by weighted Jacobi with the plate grounded (warm-started from the
    previous step; a cold start may run 2 n_max^2 sweeps), checked every 10 sweeps against the
    relative L1 change; then J = -sigma grad(phi), the enclosed axial current per radius and
    slab, the axisymmetric B_theta and the force on the melt.
    """
    from .. import kernels

    if not twin.enable_lorentz:
        return
    g.ensure_lorentz_fields()
    kernels.elec_build_sigma(g.sigma_elec, g.f_l, g.flags, twin.sigma_liquid_Sm,
                             twin.sigma_solid_Sm, g.FLAG_GAS)
    kernels.elec_clear_source(g.elec_source)
    kernels.elec_inject_arc_source(g.elec_source, arc_i, arc_j, arc_k, twin.sigma_cells,
                                   twin.welding_current_A, g.dx, g.nz)
    kernels.elec_normalize_source(g.elec_source, twin.welding_current_A, g.dx, scratch=g.scratch)
    cold_start = not getattr(twin, "_lorentz_warm", False)
    if cold_start:
        kernels.elec_init_ground(g.phi_elec, g.flags, twin.nz_solid, g.FLAG_SOLID)
        twin._lorentz_warm = True
    max_iters = max(int(twin.lorentz_jacobi_iters), 20)
    if cold_start:
        n_max = max(g.nx, g.ny, g.nz)
        max_iters = max(max_iters, 2 * n_max * n_max)
    tol = float(getattr(twin, "lorentz_jacobi_tol", 1e-4))
    check_every = 10
    converged = False
    iters_done = 0
    rel = float("inf")
    while iters_done < max_iters:
        chunk = min(check_every, max_iters - iters_done)
        for it in range(chunk):
            kernels.elec_jacobi_step(g.phi_elec, g.phi_elec_tmp, g.sigma_elec, g.elec_source,
                                     g.flags, g.dx, twin.lorentz_jacobi_omega, g.FLAG_GAS,
                                     g.FLAG_SOLID, twin.nz_solid, g.nx, g.ny, g.nz)
            if it == chunk - 1:
                # only the last sweep's change is read (the reference overwrites the others)
                kernels.elec_l1_diff_into(g.phi_elec_tmp, g.phi_elec, g.elec_l1_buf, g.scratch)
            g.phi_elec, g.phi_elec_tmp = g.phi_elec_tmp, g.phi_elec
            iters_done += 1
        res, norm = (float(v) for v in g.elec_l1_buf.to_numpy())
        rel = res / max(norm, 1e-30)
        if rel < tol:
            converged = True
            break
    if not converged:
        twin._lorentz_unconverged = getattr(twin, "_lorentz_unconverged", 0) + 1
        twin._lorentz_unconverged_streak = getattr(twin, "_lorentz_unconverged_streak", 0) + 1
        n = twin._lorentz_unconverged
        if n in (1, 10, 100) or n % 1000 == 0:
            import logging

            logging.getLogger("waam_solver_amd").warning(
                f"[weld_forces] Lorentz Jacobi hit {max_iters} iterations without converging "
                f"(rel change {rel:.2e} > {tol:.0e}); raise lorentz_jacobi_iters ({n} times)")
    else:
        twin._lorentz_unconverged_streak = 0
    kernels.elec_compute_J(g.Jx, g.Jy, g.Jz, g.phi_elec, g.sigma_elec, g.flags, g.dx, g.FLAG_GAS,
                           g.nx, g.ny, g.nz)
    bin_dr_cells = max(float(max(g.nx, g.ny)) / (2.0 * g.n_elec_bins), 0.5)
    kernels.elec_bin_axial_current(g.Jz, g.elec_rad_bins, arc_i, arc_j, g.dx, bin_dr_cells,
                                   g.n_elec_bins, g.nz)
    kernels.elec_prefix_bins(g.elec_rad_bins, g.n_elec_bins, g.nz)
    kernels.elec_B_axisymmetric(g.Bx, g.By, g.Bz, g.elec_rad_bins, arc_i, arc_j, g.dx,
                                bin_dr_cells, MU0, g.n_elec_bins, g.nz)
    kernels.apply_lorentz_JxB(g.Fx, g.Fy, g.Fz, g.Jx, g.Jy, g.Jz, g.Bx, g.By, g.Bz, g.f_l,
                              g.flags, g.dt, g.dx, twin.mat.rho, g.FLAG_GAS)
