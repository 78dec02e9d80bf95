"""One coupled physics timestep (reference solvers/coupled_step.py:17-464).

``coupled_step`` launches the HIP kernels in the reference's order, one launch per reference
kernel, on the current torch stream. Differences from the reference's host code, all of them
keeping its results:

* in-place neighbour-reading kernels (thermal advection, contact-angle BC) write a partner
  buffer that is then swapped with the field, as is ``phi_tmp`` after VOF advection (the
  reference copies it back);
* the droplet placement loop (up to ten feed_wire_surface calls with a growing footprint) runs
  in a single device launch; the host reads back the placed volume once per droplet for the
  mass ledger;
* arc_k is read back once per step, as in the reference.
"""

from __future__ import annotations

import logging

import numpy as np

from .. import kernels
from ..field import P
from ..physics import deposition, weld_forces
from ..physics.electrical_stickout import droplet_entry_temperature_K, update_ctwd

log = logging.getLogger("waam_solver_amd")


def _swap(g, a: str, b: str) -> None:
    fa, fb = getattr(g, a), getattr(g, b)
    setattr(g, a, fb)
    setattr(g, b, fa)


def _clamp_enthalpy_ceiling(twin, g) -> None:
    """Reference coupled_step.py:17-31."""
    if not twin.enable_enthalpy_cap:
        return
    if twin.use_material_tables:
        kernels.clamp_enthalpy_ceiling_variable_cp(g.H, g.flags, g.cp_rho_field, twin.H_liq,
                                                   twin.mat.T_liquidus, twin.T_vapor_cap_K,
                                                   g.FLAG_GAS)
        return
    kernels.clamp_enthalpy_ceiling_scalar(g.H, g.flags, twin.cp_rho, twin.mat.T_solidus,
                                          twin.mat.T_liquidus, twin.T_vapor_cap_K, twin.L_rho,
                                          g.FLAG_GAS)


def _refresh_properties(twin, g) -> None:
    """Per-cell properties from the material tables (reference coupled_step.py:168-182)."""
    t = twin.gpu_tables.struct
    kernels.refresh_thermal_properties(
        g.T, g.cp_rho_field, g.alpha_lu_field, g.dgamma_lu_field, g.tau_field, t.cp_T, t.cp_V,
        t.k_T, t.k_V, t.mu_T, t.mu_V, t.dgamma_T, t.dgamma_V, t.n_cp, t.n_k, t.n_mu, t.n_dgamma,
        g.mat.rho, g.dt, g.dx, t.cp_fallback, t.k_fallback, t.mu_fallback, t.dgamma_fallback,
        twin.force_scale / g.dx, twin.cp_rho, twin.alpha_lu, twin.dgamma_dT_lu, g.tau,
        twin.marangoni_scale, 1, g.flags, g.FLAG_SOLID, g.FLAG_GAS)


def _update_phase(twin, g) -> None:
    if twin.use_material_tables:
        kernels.update_phase_variable_cp(g.H, g.T, g.f_l, g.cp_rho_field, twin.L_rho,
                                         twin.mat.T_solidus, twin.mat.T_liquidus, twin.H_sol,
                                         twin.H_liq)
    else:
        kernels.update_phase(g.H, g.T, g.f_l, twin.cp_rho, twin.L_rho, twin.mat.T_solidus,
                             twin.mat.T_liquidus)


def resolve_arc_k(twin, g, arc_i: float, arc_j: float, torch_z_m: float | None = None) -> float:
    """Top metal cell in the torch column (reference coupled_step.py:34-59)."""
    i0 = max(0, min(int(arc_i), g.nx - 1))
    j0 = max(0, min(int(arc_j), g.ny - 1))
    kernels.surface_height_at(g.phi, g.flags, g.surface_k_buf, i0, j0, twin.nz_solid, g.FLAG_GAS,
                              g.nz)
    k_surface = float(g.surface_k_buf[None])
    if k_surface < 1.0:
        k_surface = float(max(1, twin.nz_solid - 1))
    if getattr(twin, "use_torch_z", False) and torch_z_m is not None:
        z_bead = max(twin.substrate_z_m, torch_z_m - twin.ctwd_m)
        k_robot = twin.nz_solid + (z_bead - twin.substrate_z_m) / g.dx - 1.0
        k_robot = max(float(twin.nz_solid) + 0.5, min(k_robot, float(g.nz) - 2.0))
        return max(k_surface, k_robot)
    return k_surface


def droplet_due(twin, g) -> bool:
    """Whether a droplet detaches on this step (reference coupled_step.py:110-117)."""
    if twin.droplet_freq <= 0:
        return False
    period = deposition.droplet_period_s(twin)
    return not (twin._step_n * g.dt - twin._last_droplet_time < period)


def droplet_plan(twin, g) -> tuple[float, float, float, float, float]:
    """The host side of a detaching droplet (reference coupled_step.py:110-124, 136): advances
    the droplet clock; (drop_mass kg, drop_r cells, drop_vol m^3, T_drop K, footprint cells)."""
    sim_time = twin._step_n * g.dt
    drop_dt = sim_time - twin._last_droplet_time
    twin._last_droplet_time = sim_time
    drop_mass = deposition.droplet_mass_for_interval_kg(twin, drop_dt)
    drop_r = deposition.droplet_radius_cells_from_mass_kg(twin, drop_mass)
    drop_vol = drop_mass / twin.mat.rho
    T_drop = droplet_entry_temperature_K(twin)
    foot_r = deposition.deposition_footprint_cells(twin, drop_r)
    return drop_mass, drop_r, drop_vol, T_drop, foot_r


def _droplet(twin, g, x_m: float, y_m: float, arc_i: float, arc_j: float, arc_k: float) -> None:
    """Droplet block (reference coupled_step.py:110-166)."""
    if not droplet_due(twin, g):
        return
    drop_mass, drop_r, drop_vol, T_drop, foot_r = droplet_plan(twin, g)
    kernels.inject_tracers(g.porosity_pos, g.porosity_active, g.tracer_head, g.max_tracers, x_m,
                           y_m, float(arc_k + 1) * g.dx, float(twin.sigma_cells * g.dx), 50,
                           seed=twin.tracer_seed)
    from .._lib import lib
    import torch

    lib().call("waam_deposit_droplet", P(g.f_src), P(g.flags), P(g.f_l), P(g.phi), P(g.H), P(g.T),
               P(g.rho), arc_i, arc_j, arc_k, float(foot_r), drop_r, float(drop_vol), T_drop,
               twin.cp_rho, twin.L_rho, 1.0, P(g.deposit_vol_buf), P(g.deposit_real_buf),
               g.dx ** 3, P(g.deposit_result), g.nx, g.ny, g.nz, g.layout, P(g.scratch),
               torch.cuda.current_stream().cuda_stream)
    placed = float(g.deposit_result[0])
    if placed < drop_vol * 0.98:
        twin._deposition_overflow += 1
        if not twin._warned_overflow and twin._deposition_overflow >= 20:
            twin._warned_overflow = True
            log.warning(f"[coupled_step] {twin._deposition_overflow} droplets deposited less "
                        "volume than the wire feed supplies; mass_balance_ratio will drift")
    twin._deposited_volume_m3 += placed
    twin._n_droplets_fired += 1
    weld_forces.apply_droplet_impact(twin, g, arc_i, arc_j, arc_k, drop_r, drop_mass)


def coupled_step(twin, torch_x_m: float, torch_y_m: float, is_welding: bool,
                 torch_z_m: float | None = None) -> None:
    g = twin.grid
    tables = twin.use_material_tables
    n = g.nx, g.ny, g.nz
    arc_i = torch_x_m / g.dx
    arc_j = torch_y_m / g.dx
    arc_k = resolve_arc_k(twin, g, arc_i, arc_j, torch_z_m)
    twin._last_arc_ijk = (float(arc_i), float(arc_j), float(arc_k))

    kernels.clear_forces(g.Fx, g.Fy, g.Fz)
    if twin.enable_ctwd:
        update_ctwd(twin, g)
    current_pressure = weld_forces.arc_pressure_peak_pa(twin)
    pressure_sigma_cells = weld_forces.arc_pressure_sigma_cells(twin)

    if is_welding:
        twin.arc_source.inject(twin, g, arc_i, arc_j, arc_k)
        _clamp_enthalpy_ceiling(twin, g)
        _droplet(twin, g, torch_x_m, torch_y_m, arc_i, arc_j, arc_k)

    # thermal advection-diffusion, Jacobi: H_tmp <- H + dH(H); swap
    if tables:
        _refresh_properties(twin, g)
        kernels.advect_diffuse_temperature_variable_into(g.H, g.H_tmp, g.T, g.ux, g.uy, g.uz,
                                                         g.flags, g.alpha_lu_field,
                                                         g.cp_rho_field, 1.0, *n)
    else:
        kernels.advect_diffuse_temperature_into(g.H, g.H_tmp, g.T, g.ux, g.uy, g.uz, g.flags,
                                                twin.alpha_lu, 1.0, twin.cp_rho, *n)
    _swap(g, "H", "H_tmp")
    if twin.enable_heat_loss:
        conv, rad = (1 if twin.enable_convection else 0), (1 if twin.enable_radiation else 0)
        if tables:
            kernels.apply_thermal_boundary_losses_variable(
                g.H, g.T, g.flags, g.cp_rho_field, twin.T_amb, twin.h_conv, twin.eps_rad, conv,
                rad, g.dt, g.dx, twin.sigma_sb, g.FLAG_SOLID, g.FLAG_GAS, *n)
            kernels.clamp_enthalpy_floor(g.H, g.cp_rho_field, g.flags, twin.T_amb, g.FLAG_GAS)
        else:
            kernels.apply_thermal_boundary_losses(
                g.H, g.T, g.flags, twin.T_amb, twin.h_conv, twin.eps_rad, conv, rad, twin.cp_rho,
                g.dt, g.dx, twin.sigma_sb, g.FLAG_SOLID, g.FLAG_GAS, *n)
            kernels.clamp_enthalpy_floor_scalar(g.H, twin.cp_rho, g.flags, twin.T_amb,
                                                g.FLAG_GAS)
    _update_phase(twin, g)
    _clamp_enthalpy_ceiling(twin, g)
    _update_phase(twin, g)
    kernels.update_T_max(g.T, g.T_max, g.flags, g.FLAG_GAS)
    kernels.update_cooling_rate(g.T, g.T_prev, g.dT_dt, g.flags, g.dt, g.FLAG_GAS)

    if twin.enable_vof:
        g.ensure_vof_buffers()
        kernels.advect_phi(g.phi_tmp, g.phi, g.ux, g.uy, g.uz, g.flags, g.FLAG_SOLID, g.FLAG_GAS,
                           *n)
        _swap(g, "phi", "phi_tmp")
        kernels.reinitialize_phi(g.phi, g.flags, g.FLAG_SOLID, g.FLAG_GAS, g.FLAG_FLUID)
        if twin.enable_wetting:
            kernels.apply_contact_angle_phi_bc_into(g.phi, g.phi_tmp, g.flags, twin.theta_rad, *n)
            _swap(g, "phi", "phi_tmp")
        kernels.update_flags_from_phi(g.phi, g.f_l, g.flags, twin.nz_solid, g.FLAG_FLUID,
                                      g.FLAG_SOLID, g.FLAG_GAS, g.FLAG_IFACE)
        kernels.sync_phi_liquid_fraction(g.phi, g.f_l, g.flags, g.FLAG_GAS, g.FLAG_SOLID)

    if twin.enable_csf_tension:
        kernels.compute_csf_tension(g.phi, g.flags, g.Fx, g.Fy, g.Fz, twin.gamma_lu, g.FLAG_SOLID,
                                    g.FLAG_GAS, *n, int(twin.enable_wetting), twin.theta_rad)
    if tables:
        kernels.compute_marangoni_force_variable(g.T, g.phi, g.f_l, g.Fx, g.Fy, g.Fz, g.flags,
                                                 g.dgamma_lu_field, g.FLAG_SOLID, g.FLAG_GAS, *n)
    else:
        kernels.compute_marangoni_force(g.T, g.phi, g.f_l, g.Fx, g.Fy, g.Fz, g.flags,
                                        twin.dgamma_dT_lu, g.dx, g.FLAG_SOLID, g.FLAG_GAS, *n)
    if is_welding and twin.enable_gas_shear:
        weld_forces.apply_gas_shear(twin, g, arc_i, arc_j, arc_k)
    if is_welding:
        kernels.apply_arc_pressure(g.Fz, g.flags, g.phi, arc_i, arc_j, arc_k,
                                   pressure_sigma_cells, current_pressure, g.dt, g.dx,
                                   twin.mat.rho, g.FLAG_SOLID, g.FLAG_GAS)
        weld_forces.apply_recoil(twin, g, arc_i, arc_j, arc_k)
    if twin.enable_hydrostatic_gravity:
        kernels.add_hydrostatic_gravity(g.Fz, g.f_l, g.flags, 1.0, twin.g_lu, g.FLAG_SOLID,
                                        g.FLAG_GAS)
    kernels.add_buoyancy(g.T, g.Fz, g.f_l, g.flags, twin.g_lu, twin.beta_T, twin.mat.T_liquidus,
                         1.0, g.FLAG_SOLID, g.FLAG_GAS)
    if is_welding and twin.enable_lorentz:
        weld_forces.solve_lorentz(twin, g, arc_i, arc_j, arc_k)
    u_cap = float(getattr(twin, "u_mach_limit_lu", 0.08))
    F_cap = float(getattr(twin, "force_limit_lu", 0.05))
    if F_cap > 0.0:
        kernels.clamp_body_force_magnitude(g.Fx, g.Fy, g.Fz, g.flags, F_cap, g.FLAG_SOLID,
                                           g.FLAG_GAS)
    if tables and twin.use_variable_tau:
        kernels.collide_srt_variable_tau(g.f_src, g.f_dst, g.rho, g.ux, g.uy, g.uz, g.Fx, g.Fy,
                                         g.Fz, g.f_l, g.flags, g.tau_field, twin.C_darcy,
                                         g.FLAG_SOLID, g.FLAG_GAS, *n)
    elif twin.use_srt:
        kernels.collide_srt(g.f_src, g.f_dst, g.rho, g.ux, g.uy, g.uz, g.Fx, g.Fy, g.Fz, g.f_l,
                            g.flags, g.tau, twin.omega, 1.0, twin.C_darcy, g.FLAG_SOLID,
                            g.FLAG_GAS, *n)
    else:
        kernels.collide_mrt(g.f_src, g.f_dst, g.rho, g.ux, g.uy, g.uz, g.Fx, g.Fy, g.Fz, g.f_l,
                            g.flags, None, None, None, None, None, twin.omega, twin.omega_bulk,
                            twin.C_darcy, g.FLAG_SOLID, g.FLAG_GAS, *n)
    if u_cap > 0.0:
        kernels.clamp_velocity_mach(g.ux, g.uy, g.uz, g.flags, u_cap, g.FLAG_SOLID, g.FLAG_GAS)
    kernels.stream(g.f_dst, g.f_src, g.flags, g.FLAG_SOLID, g.FLAG_GAS, *n)
    kernels.advect_tracers(g.porosity_pos, g.porosity_active, g.ux, g.uy, g.uz, g.f_l, g.flags,
                           g.dx, g.dt, g.max_tracers, g.FLAG_SOLID, g.FLAG_GAS)
    kernels.update_time_above_T(g.T, g.flags, g.time_above_800_s, g.time_above_1100_s,
                                g.time_above_solidus_s, g.dt, 800.0 + 273.15, 1100.0 + 273.15,
                                twin.mat.T_solidus, g.FLAG_GAS)
    if twin.snapshot_forces:
        kernels.snapshot_forces(g.Fx, g.Fy, g.Fz, g.Fx_snap, g.Fy_snap, g.Fz_snap)
    if twin.enable_substrate_growth or twin.enable_bead_freeze:
        if tables:
            kernels.remelt_hot_solid(g.T, g.H, g.f_l, g.phi, g.flags, g.cp_rho_field,
                                     twin.mat.T_solidus, twin.mat.T_liquidus, twin.L_rho,
                                     g.FLAG_SOLID, g.FLAG_FLUID)
        else:
            kernels.remelt_hot_solid_scalar(g.T, g.H, g.f_l, g.phi, g.flags, twin.cp_rho,
                                            twin.mat.T_solidus, twin.mat.T_liquidus, twin.L_rho,
                                            g.FLAG_SOLID, g.FLAG_FLUID)
        if is_welding and twin.enable_bead_freeze:
            dir_x, dir_y, dir_z = getattr(twin, "_torch_dir_xyz", (1.0, 0.0, 0.0))
            lookback = max(2.0, twin.trailing_solidify_lookback_mm / (g.dx * 1000.0))
            T_freeze = twin.mat.T_liquidus + twin.trailing_solidify_temp_margin_K
            if tables:
                kernels.solidify_trailing_pool(
                    g.T, g.H, g.f_l, g.phi, g.flags, g.ux, g.uy, g.uz, g.cp_rho_field, arc_i,
                    arc_j, arc_k, dir_x, dir_y, dir_z, lookback, T_freeze, twin.mat.T_solidus,
                    g.FLAG_SOLID, g.FLAG_FLUID, g.FLAG_GAS)
            else:
                kernels.solidify_trailing_pool_scalar(
                    g.T, g.H, g.f_l, g.phi, g.flags, g.ux, g.uy, g.uz, twin.cp_rho, arc_i, arc_j,
                    arc_k, dir_x, dir_y, dir_z, lookback, T_freeze, twin.mat.T_solidus,
                    g.FLAG_SOLID, g.FLAG_FLUID, g.FLAG_GAS)
        kernels.solidify_cooled_metal(g.T, g.f_l, g.phi, g.flags, g.ux, g.uy, g.uz,
                                      twin.mat.T_solidus,
                                      twin.enable_bead_freeze or twin.enable_substrate_growth,
                                      g.FLAG_SOLID, g.FLAG_FLUID, g.FLAG_GAS)
    g.swap_buffers()
    twin._step_n += 1
