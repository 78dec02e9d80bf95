"""Positional kernel API of the reference (kernels.py / cumulant_kernel.py) on the HIP library.

Every function keeps the reference kernel's name and positional parameters — fields first,
then scalars, FLAG_* constants and grid sizes — and launches the matching libwaam_hip entry
point on the current torch stream. Fields are DeviceFields (or contiguous CUDA tensors). FLAG_*
arguments are checked against the fixed flag values of the library. Sizes are checked against
the tensors before any launch.

Two reference kernels update H or phi in place while reading neighbouring values of the same
field (advect_diffuse_temperature, apply_contact_angle_phi_bc); here they read a copy of the
pre-kernel field (Jacobi order), which is the one deterministic order a parallel GPU kernel has.
"""

from __future__ import annotations

import math

import torch

from ._lib import FLAG_FLUID, FLAG_GAS, FLAG_IFACE, FLAG_SOLID, WaamError, lib
from .field import DeviceField, P, dims_of, layout_of, note_write, tensor_of

_FLAGS = {"FLAG_FLUID": FLAG_FLUID, "FLAG_SOLID": FLAG_SOLID, "FLAG_GAS": FLAG_GAS,
          "FLAG_IFACE": FLAG_IFACE}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _flags_ok(**kw) -> None:
    for k, v in kw.items():
        if int(v) != _FLAGS[k]:
            raise ValueError(f"{k}={v} differs from the library's fixed value {_FLAGS[k]}")


def _need(n: int, *fields) -> None:
    for f in fields:
        if f is None:
            continue
        t = tensor_of(f)
        if t.numel() < n:
            raise ValueError(f"field of {t.numel()} elements is smaller than the {n} the call needs")


def _need_pop(n: int, *fields) -> None:
    _need(19 * n, *fields)


def _n(x) -> int:
    return tensor_of(x).numel()


def bind_velocity_set(grid) -> None:
    """The velocity set is compiled into the library (kept for API compatibility)."""
    return None


def _call(name, *args):
    note_write()  # any per-kernel entry may write fields the fused step tracks
    lib().call(name, *args)


def _call_ro(name, *args):
    """An entry that reads fields and writes only its own output buffer."""
    lib().call(name, *args)


def _scratch_of(*fields):
    """Workspace tensor for reducing entries: the grid's when reachable, else a fresh one."""
    return torch.zeros(lib().scratch_bytes(), dtype=torch.uint8, device=tensor_of(fields[0]).device)


# ------------------------------------------------------------------ init / forces reset
def init_grid(f, rho, ux, uy, uz, T, H, f_l, phi, flags, T_ambient, rho0, cp_rho, nz_solid,
              plate_i0, plate_i1, plate_j0, plate_j1, FLAG_FLUID, FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_FLUID=FLAG_FLUID, FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(rho)
    n = nx * ny * nz
    _need_pop(n, f)
    _need(n, rho, ux, uy, uz, T, H, f_l, phi, flags)
    _call("waam_init_grid", P(f), P(rho), P(ux), P(uy), P(uz), P(T), P(H), P(f_l), P(phi),
          P(flags), T_ambient, rho0, cp_rho, nz_solid, plate_i0, plate_i1, plate_j0, plate_j1,
          nx, ny, nz, layout_of(f), _stream())


def init_aux_fields(T_max, T_prev, dT_dt, time_above_800, time_above_1100, time_above_solidus,
                    Fx_snap, Fy_snap, Fz_snap, Fx, Fy, Fz, T_ambient, porosity_active,
                    tracer_head, max_tracers):
    for fld in (T_max, T_prev):
        tensor_of(fld).fill_(float(T_ambient))
    for fld in (dT_dt, time_above_800, time_above_1100, time_above_solidus, Fx_snap, Fy_snap,
                Fz_snap, Fx, Fy, Fz):
        tensor_of(fld).zero_()
    tensor_of(porosity_active).zero_()
    tensor_of(tracer_head).zero_()


def clear_forces(Fx, Fy, Fz):
    for fld in (Fx, Fy, Fz):
        tensor_of(fld).zero_()


# ------------------------------------------------------------------ arc heat
def inject_arc_heat(H, flags, phi, f_l, norm_buf, arc_i, arc_j, arc_k, Q_w, sigma, dt, dx3, eta,
                    penetration_cells, enable_surface_weight, FLAG_SOLID, FLAG_GAS, scratch=None):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(H)
    _need(nx * ny * nz, flags, f_l)
    sc = scratch if scratch is not None else _scratch_of(H)
    _call("waam_inject_arc_heat", P(H), P(flags), P(f_l), P(norm_buf), arc_i, arc_j, arc_k, Q_w,
          sigma, dt, dx3, eta, penetration_cells, int(enable_surface_weight), nx, ny, nz, layout_of(H), P(sc),
          _stream())


def inject_goldak_heat(H, flags, phi, f_l, norm_buf, arc_i, arc_j, arc_k, Q_w, dt, dx3, eta,
                       travel_sign, ff, fr, a_front, a_rear, b_axis, c_axis, penetration_cells,
                       enable_surface_weight, FLAG_SOLID, FLAG_GAS, scratch=None):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(H)
    _need(nx * ny * nz, flags, f_l)
    sc = scratch if scratch is not None else _scratch_of(H)
    _call("waam_inject_goldak_heat", P(H), P(flags), P(f_l), P(norm_buf), arc_i, arc_j, arc_k,
          Q_w, dt, dx3, eta, travel_sign, ff, fr, a_front, a_rear, b_axis, c_axis,
          penetration_cells, int(enable_surface_weight), nx, ny, nz, layout_of(H), P(sc), _stream())


def surface_height_at(phi, flags, out, i0, j0, nz_fallback, FLAG_GAS, nz):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    nx, ny, nz_ = dims_of(phi)
    if nz != nz_:
        raise ValueError("nz does not match phi")
    _call("waam_surface_height_at", P(phi), P(flags), P(out), int(i0), int(j0), int(nz_fallback),
          nx, ny, nz, layout_of(phi), _stream())


# ------------------------------------------------------------------ thermal
def clamp_enthalpy_ceiling_scalar(H, flags, cp_rho, T_solidus, T_liquidus, T_vapor_cap, L_rho,
                                  FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(H), flags)
    _call("waam_clamp_enthalpy_ceiling_scalar", P(H), P(flags), cp_rho, T_solidus, T_liquidus,
          T_vapor_cap, L_rho, _n(H), _stream())


def clamp_enthalpy_floor_scalar(H, cp_rho, flags, T_floor, FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(H), flags)
    _call("waam_clamp_enthalpy_floor_scalar", P(H), cp_rho, P(flags), T_floor, _n(H), _stream())


def advect_diffuse_temperature_into(H_in, H_out, T, ux, uy, uz, flags, alpha_lu, dt, cp_rho, nx,
                                    ny, nz):
    _need(nx * ny * nz, H_in, H_out, T, ux, uy, uz, flags)
    _call("waam_advect_diffuse_temperature", P(H_in), P(H_out), P(T), P(ux), P(uy), P(uz),
          P(flags), alpha_lu, dt, cp_rho, nx, ny, nz, layout_of(H_in), _stream())


def advect_diffuse_temperature(H, T, ux, uy, uz, flags, alpha_lu, dt, FLAG_SOLID, FLAG_GAS,
                               cp_rho, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    H_in = tensor_of(H).clone()
    advect_diffuse_temperature_into(H_in, H, T, ux, uy, uz, flags, alpha_lu, dt, cp_rho, nx, ny,
                                    nz)


def apply_thermal_boundary_losses(H, T, flags, T_amb, h_conv, eps_rad, enable_conv, enable_rad,
                                  cp_rho, dt, dx, sigma_sb, FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, H, T, flags)
    _call("waam_apply_thermal_boundary_losses", P(H), P(T), P(flags), T_amb, h_conv, eps_rad,
          int(enable_conv), int(enable_rad), cp_rho, dt, dx, sigma_sb, nx, ny, nz, layout_of(H), _stream())


def update_phase(H, T, f_l, cp_rho, L_rho, T_solidus, T_liquidus):
    _need(_n(H), T, f_l)
    _call("waam_update_phase", P(H), P(T), P(f_l), cp_rho, L_rho, T_solidus, T_liquidus, _n(H),
          _stream())


def update_T_max(T, T_max, flags, FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(T), T_max, flags)
    _call("waam_update_T_max", P(T), P(T_max), P(flags), _n(T), _stream())


def update_cooling_rate(T, T_prev, dT_dt, flags, dt_phys, FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(T), T_prev, dT_dt, flags)
    _call("waam_update_cooling_rate", P(T), P(T_prev), P(dT_dt), P(flags), dt_phys, _n(T),
          _stream())


def update_time_above_T(T, flags, time_above_800, time_above_1100, time_above_solidus, dt,
                        T_800, T_1100, T_solidus, FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(T), flags, time_above_800, time_above_1100, time_above_solidus)
    _call("waam_update_time_above_T", P(T), P(flags), P(time_above_800), P(time_above_1100),
          P(time_above_solidus), dt, T_800, T_1100, T_solidus, _n(T), _stream())


def sync_T_from_H(H, T, cp_rho):
    """T = H / (rho cp) (reference kernels.py:3119-3122; a verification helper)."""
    tensor_of(T).copy_(tensor_of(H) / torch.tensor(cp_rho, dtype=torch.float32))


# ------------------------------------------------------------------ VOF
def advect_phi(phi_dst, phi_src, ux, uy, uz, flags, FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, phi_dst, phi_src, ux, uy, uz, flags)
    _call("waam_advect_phi", P(phi_dst), P(phi_src), P(ux), P(uy), P(uz), P(flags), nx, ny, nz, layout_of(phi_dst),
          _stream())


def reinitialize_phi(phi, flags, FLAG_SOLID, FLAG_GAS, FLAG_FLUID):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS, FLAG_FLUID=FLAG_FLUID)
    _need(_n(phi), flags)
    _call("waam_reinitialize_phi", P(phi), P(flags), _n(phi), _stream())


def contact_angle_cos_sin(theta_rad: float) -> tuple[float, float]:
    """cos/sin of the fp32-rounded angle, rounded to fp32 (the reference's ti.cos/ti.sin)."""
    import numpy as np

    t = float(np.float32(theta_rad))
    return float(np.float32(math.cos(t))), float(np.float32(math.sin(t)))


def apply_contact_angle_phi_bc_into(phi_in, phi_out, flags, theta_rad, nx, ny, nz):
    _need(nx * ny * nz, phi_in, phi_out, flags)
    c, s = contact_angle_cos_sin(theta_rad)
    _call("waam_apply_contact_angle_phi_bc", P(phi_in), P(phi_out), P(flags), c, s, nx, ny, nz, layout_of(phi_in),
          _stream())


def apply_contact_angle_phi_bc(phi, flags, theta_rad, FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    p_in = tensor_of(phi).clone()
    apply_contact_angle_phi_bc_into(p_in, phi, flags, theta_rad, nx, ny, nz)


def update_flags_from_phi(phi, f_l, flags, nz_solid, FLAG_FLUID, FLAG_SOLID, FLAG_GAS,
                          FLAG_IFACE):
    _flags_ok(FLAG_FLUID=FLAG_FLUID, FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS,
              FLAG_IFACE=FLAG_IFACE)
    nx, ny, nz = dims_of(phi)
    _need(nx * ny * nz, f_l, flags)
    _call("waam_update_flags_from_phi", P(phi), P(f_l), P(flags), int(nz_solid), nx, ny, nz, layout_of(phi),
          _stream())


def sync_phi_liquid_fraction(phi, f_l, flags, FLAG_GAS, FLAG_SOLID):
    _flags_ok(FLAG_GAS=FLAG_GAS, FLAG_SOLID=FLAG_SOLID)
    _need(_n(phi), f_l, flags)
    _call("waam_sync_phi_liquid_fraction", P(phi), P(f_l), P(flags), _n(phi), _stream())


def solidify_cooled_metal(T, f_l, phi, flags, ux, uy, uz, T_solidus, zero_velocity, FLAG_SOLID,
                          FLAG_FLUID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_FLUID=FLAG_FLUID, FLAG_GAS=FLAG_GAS)
    _need(_n(T), f_l, phi, flags, ux, uy, uz)
    _call("waam_solidify_cooled_metal", P(T), P(f_l), P(phi), P(flags), P(ux), P(uy), P(uz),
          T_solidus, 1 if zero_velocity else 0, _n(T), _stream())


def remelt_hot_solid_scalar(T, H, f_l, phi, flags, cp_rho, T_solidus, T_liquidus, L_rho,
                            FLAG_SOLID, FLAG_FLUID):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_FLUID=FLAG_FLUID)
    _need(_n(T), H, f_l, phi, flags)
    _call("waam_remelt_hot_solid_scalar", P(T), P(H), P(f_l), P(phi), P(flags), cp_rho, T_solidus,
          T_liquidus, L_rho, _n(T), _stream())


def solidify_trailing_pool_scalar(T, H, f_l, phi, flags, ux, uy, uz, cp_rho, arc_i, arc_j, arc_k,
                                  dir_x, dir_y, dir_z, lookback_cells, T_freeze, T_solidus,
                                  FLAG_SOLID, FLAG_FLUID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_FLUID=FLAG_FLUID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(T)
    _need(nx * ny * nz, H, f_l, phi, flags, ux, uy, uz)
    _call("waam_solidify_trailing_pool_scalar", P(T), P(H), P(f_l), P(phi), P(flags), P(ux),
          P(uy), P(uz), cp_rho, arc_i, arc_j, arc_k, dir_x, dir_y, dir_z, lookback_cells,
          T_freeze, T_solidus, nx, ny, nz, layout_of(T), _stream())


# ------------------------------------------------------------------ forces
def compute_csf_tension(phi, flags, Fx, Fy, Fz, gamma_lu, FLAG_SOLID, FLAG_GAS, nx, ny, nz,
                        enable_wetting=0, theta_rad=0.0):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, phi, flags, Fx, Fy, Fz)
    c, s = contact_angle_cos_sin(theta_rad)
    _call("waam_compute_csf_tension", P(phi), P(flags), P(Fx), P(Fy), P(Fz), gamma_lu,
          int(enable_wetting), c, s, nx, ny, nz, layout_of(phi), _stream())


compute_csf_wetting = compute_csf_tension


def compute_curvature_field(phi, flags, kappa_out, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, phi, flags, kappa_out)
    _call("waam_compute_curvature_field", P(phi), P(flags), P(kappa_out), nx, ny, nz, layout_of(phi), _stream())


def compute_marangoni_force(T, phi, f_l, Fx, Fy, Fz, flags, dgamma_dT, dx, FLAG_SOLID, FLAG_GAS,
                            nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, T, phi, f_l, Fx, Fy, Fz, flags)
    _call("waam_compute_marangoni_force", P(T), P(phi), P(f_l), P(Fx), P(Fy), P(Fz), P(flags),
          dgamma_dT, nx, ny, nz, layout_of(T), _stream())


def add_buoyancy(T, Fz, f_l, flags, g_lu, beta, T_ref, rho_ref, FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(_n(T), Fz, f_l, flags)
    _call("waam_add_buoyancy", P(T), P(Fz), P(f_l), P(flags), g_lu, beta, T_ref, rho_ref, _n(T),
          _stream())


def add_hydrostatic_gravity(Fz, f_l, flags, rho_ref, g_lu, FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(_n(Fz), f_l, flags)
    _call("waam_add_hydrostatic_gravity", P(Fz), P(f_l), P(flags), rho_ref, g_lu, _n(Fz),
          _stream())


def apply_arc_pressure(Fz, flags, phi, arc_i, arc_j, arc_k, sigma, pressure_pa, dt, dx, rho_ref,
                       FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(Fz)
    _need(nx * ny * nz, flags, phi)
    _call("waam_apply_arc_pressure", P(Fz), P(flags), P(phi), arc_i, arc_j, arc_k, sigma,
          pressure_pa, dt, dx, rho_ref, nx, ny, nz, layout_of(Fz), _stream())


def clamp_body_force_magnitude(Fx, Fy, Fz, flags, F_max_lu, FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(_n(Fx), Fy, Fz, flags)
    _call("waam_clamp_body_force_magnitude", P(Fx), P(Fy), P(Fz), P(flags), F_max_lu, _n(Fx),
          _stream())


def apply_droplet_impact_pressure(Fz, flags, phi, f_l, arc_i, arc_j, arc_k, drop_radius,
                                  impact_pa, dt, dx, rho_ref, FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(Fz)
    _need(nx * ny * nz, flags, phi, f_l)
    _call("waam_apply_droplet_impact_pressure", P(Fz), P(flags), P(phi), P(f_l), arc_i, arc_j,
          arc_k, drop_radius, impact_pa, dt, dx, rho_ref, nx, ny, nz, layout_of(Fz), _stream())


def snapshot_forces(Fx, Fy, Fz, Fx_snap, Fy_snap, Fz_snap):
    for a, b in ((Fx, Fx_snap), (Fy, Fy_snap), (Fz, Fz_snap)):
        tensor_of(b).copy_(tensor_of(a))


# ------------------------------------------------------------------ LBM
def collide_mrt(f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags, EX, EY, EZ, W, OPP,
                omega_s, omega_b, C_darcy, FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    n = nx * ny * nz
    _need_pop(n, f_src, f_dst)
    _need(n, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags)
    _call("waam_collide_mrt", P(f_src), P(f_dst), P(rho), P(ux), P(uy), P(uz), P(Fx), P(Fy),
          P(Fz), P(f_l), P(flags), omega_s, omega_b, C_darcy, nx, ny, nz, layout_of(f_src), _stream())


def collide_srt(f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags, tau, omega, dt_lu, C_darcy,
                FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    n = nx * ny * nz
    _need_pop(n, f_src, f_dst)
    _need(n, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags)
    _call("waam_collide_srt", P(f_src), P(f_dst), P(rho), P(ux), P(uy), P(uz), P(Fx), P(Fy),
          P(Fz), P(f_l), P(flags), tau, omega, dt_lu, C_darcy, nx, ny, nz, layout_of(f_src), _stream())


def collide_srt_variable_tau(f_src, f_dst, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags, tau_field,
                             C_darcy, FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    n = nx * ny * nz
    _need_pop(n, f_src, f_dst)
    _need(n, rho, ux, uy, uz, Fx, Fy, Fz, f_l, flags, tau_field)
    _call("waam_collide_srt_variable_tau", P(f_src), P(f_dst), P(rho), P(ux), P(uy), P(uz), P(Fx),
          P(Fy), P(Fz), P(f_l), P(flags), P(tau_field), C_darcy, nx, ny, nz, layout_of(f_src), _stream())


def stream(f_src, f_dst, flags, FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    n = nx * ny * nz
    _need_pop(n, f_src, f_dst)
    _need(n, flags)
    _call("waam_stream", P(f_src), P(f_dst), P(flags), nx, ny, nz, layout_of(f_src), _stream())


def clamp_velocity_mach(ux, uy, uz, flags, u_max_lu, FLAG_SOLID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(_n(ux), uy, uz, flags)
    _call("waam_clamp_velocity_mach", P(ux), P(uy), P(uz), P(flags), u_max_lu, _n(ux), _stream())


# ------------------------------------------------------------------ deposition, tracers
def feed_wire_surface(f_src, flags, f_l, phi, H, T, rho, arc_i, arc_j, arc_k, footprint_r,
                      droplet_radius, target_vol, T_drop, cp_rho, L_rho, rho0, vol_acc, real_acc,
                      cell_vol, FLAG_GAS, FLAG_FLUID, FLAG_SOLID, nx, ny, nz, scratch=None):
    _flags_ok(FLAG_GAS=FLAG_GAS, FLAG_FLUID=FLAG_FLUID, FLAG_SOLID=FLAG_SOLID)
    n = nx * ny * nz
    _need_pop(n, f_src)
    _need(n, flags, f_l, phi, H, T, rho)
    sc = scratch if scratch is not None else _scratch_of(H)
    _call("waam_feed_wire_surface", P(f_src), P(flags), P(f_l), P(phi), P(H), P(T), P(rho), arc_i,
          arc_j, arc_k, footprint_r, droplet_radius, target_vol, T_drop, cp_rho, L_rho, rho0,
          P(vol_acc), P(real_acc), cell_vol, nx, ny, nz, layout_of(f_src), P(sc), _stream())


def feed_wire_momentum_impact(f_src, rho, ux, uy, uz, flags, f_l, arc_i, arc_j, arc_k,
                              droplet_radius, vx_lu, vy_lu, vz_lu, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    n = nx * ny * nz
    _need_pop(n, f_src)
    _need(n, rho, ux, uy, uz, flags, f_l)
    _call("waam_feed_wire_momentum_impact", P(f_src), P(rho), P(ux), P(uy), P(uz), P(flags),
          P(f_l), arc_i, arc_j, arc_k, droplet_radius, vx_lu, vy_lu, vz_lu, nx, ny, nz, layout_of(f_src), _stream())


def inject_tracers(pos, active, head, max_tracers, arc_x, arc_y, arc_z, sigma_m, spawn_count,
                   seed=0):
    if tensor_of(pos).numel() < 3 * max_tracers or tensor_of(active).numel() < max_tracers:
        raise ValueError("tracer buffers smaller than max_tracers")
    _call("waam_inject_tracers", P(pos), P(active), P(head), int(max_tracers), arc_x, arc_y,
          arc_z, sigma_m, int(spawn_count), int(seed) & 0xFFFFFFFF, _stream())


def advect_tracers(pos, active, ux, uy, uz, f_l, flags, dx, dt, max_tracers, FLAG_SOLID,
                   FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(ux)
    if tensor_of(pos).numel() < 3 * max_tracers or tensor_of(active).numel() < max_tracers:
        raise ValueError("tracer buffers smaller than max_tracers")
    _call("waam_advect_tracers", P(pos), P(active), P(ux), P(uy), P(uz), P(f_l), P(flags), dx, dt,
          int(max_tracers), nx, ny, nz, layout_of(pos), _stream())


# ------------------------------------------------------------------ telemetry
def telemetry_pool_reduce_into(T, f_l, phi, dT_dt, ux, uy, uz, out, scratch):
    nx, ny, nz = dims_of(T)
    _need(nx * ny * nz, f_l, phi, dT_dt, ux, uy, uz)
    _call_ro("waam_telemetry_pool_reduce", P(T), P(f_l), P(phi), P(dT_dt), P(ux), P(uy), P(uz),
          P(out), nx, ny, nz, layout_of(T), P(scratch), _stream())


# ------------------------------------------------------------------ material tables
def _host_floats(x) -> list[float]:
    if isinstance(x, DeviceField):
        return [float(v) for v in x.to_numpy().reshape(-1)]
    if isinstance(x, torch.Tensor):
        return [float(v) for v in x.detach().cpu().reshape(-1)]
    return [float(v) for v in x]


def _host_int(x) -> int:
    if isinstance(x, DeviceField):
        return int(x[None])
    return int(x)


def refresh_thermal_properties(T, cp_rho_field, alpha_lu_field, dgamma_lu_field, tau_field, cp_T,
                               cp_V, k_T, k_V, mu_T, mu_V, dgamma_T, dgamma_V, n_cp, n_k, n_mu,
                               n_dgamma, rho, dt, dx, cp_fallback, k_fallback, mu_fallback,
                               dgamma_fallback, force_scale, cp_rho_ref, alpha_lu_ref,
                               dgamma_lu_ref, tau_ref, marangoni_scale, use_tables, flags,
                               FLAG_SOLID, FLAG_GAS):
    """Per-cell cp_rho, alpha_lu, dgamma_lu and tau from the knot tables (reference
    kernels.py:1710-1772). The knot arrays may be device fields, tensors or host sequences of
    up to 8 values."""
    from .gpu_tables import MAX_KNOTS, MaterialTablesC, PropertyRefsC

    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    n = _n(T)
    _need(n, cp_rho_field, alpha_lu_field, dgamma_lu_field, tau_field, flags)
    tb = MaterialTablesC()
    for prop, t_arr, v_arr, cnt in (("cp", cp_T, cp_V, n_cp), ("k", k_T, k_V, n_k),
                                    ("mu", mu_T, mu_V, n_mu), ("dgamma", dgamma_T, dgamma_V, n_dgamma)):
        t = (_host_floats(t_arr) + [0.0] * MAX_KNOTS)[:MAX_KNOTS]
        v = (_host_floats(v_arr) + [0.0] * MAX_KNOTS)[:MAX_KNOTS]
        getattr(tb, f"{prop}_T")[:] = t
        getattr(tb, f"{prop}_V")[:] = v
        setattr(tb, f"n_{prop}", _host_int(cnt))
    tb.cp_fallback, tb.k_fallback = cp_fallback, k_fallback
    tb.mu_fallback, tb.dgamma_fallback = mu_fallback, dgamma_fallback
    rf = PropertyRefsC(rho, dt, dx, force_scale, cp_rho_ref, alpha_lu_ref, dgamma_lu_ref, tau_ref,
                       marangoni_scale)
    import ctypes

    _call("waam_refresh_thermal_properties", P(T), P(cp_rho_field), P(alpha_lu_field),
          P(dgamma_lu_field), P(tau_field), ctypes.addressof(tb), ctypes.addressof(rf),
          int(use_tables), P(flags), n, _stream())


def advect_diffuse_temperature_variable_into(H_in, H_out, T, ux, uy, uz, flags, alpha_lu_field,
                                             cp_rho_field, dt, nx, ny, nz):
    _need(nx * ny * nz, H_in, H_out, T, ux, uy, uz, flags, alpha_lu_field, cp_rho_field)
    _call("waam_advect_diffuse_temperature_variable", P(H_in), P(H_out), P(T), P(ux), P(uy),
          P(uz), P(flags), P(alpha_lu_field), P(cp_rho_field), dt, nx, ny, nz, layout_of(H_out),
          _stream())


def advect_diffuse_temperature_variable(H, T, ux, uy, uz, flags, alpha_lu_field, cp_rho_field, dt,
                                        FLAG_SOLID, FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    H_in = tensor_of(H).clone()
    advect_diffuse_temperature_variable_into(H_in, H, T, ux, uy, uz, flags, alpha_lu_field,
                                             cp_rho_field, dt, nx, ny, nz)


def update_phase_variable_cp(H, T, f_l, cp_rho_field, L_rho, T_solidus, T_liquidus, H_sol, H_liq):
    _need(_n(H), T, f_l, cp_rho_field)
    _call("waam_update_phase_variable_cp", P(H), P(T), P(f_l), P(cp_rho_field), L_rho, T_solidus,
          T_liquidus, H_sol, H_liq, _n(H), _stream())


def apply_thermal_boundary_losses_variable(H, T, flags, cp_rho_field, T_amb, h_conv, eps_rad,
                                           enable_conv, enable_rad, dt, dx, sigma_sb, FLAG_SOLID,
                                           FLAG_GAS, nx, ny, nz):
    """Reference kernels.py:1866-1918: the loss does not depend on cp_rho_field."""
    apply_thermal_boundary_losses(H, T, flags, T_amb, h_conv, eps_rad, enable_conv, enable_rad,
                                  0.0, dt, dx, sigma_sb, FLAG_SOLID, FLAG_GAS, nx, ny, nz)


def clamp_enthalpy_floor(H, cp_rho_field, flags, T_floor, FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(H), cp_rho_field, flags)
    _call("waam_clamp_enthalpy_floor", P(H), P(cp_rho_field), P(flags), T_floor, _n(H), _stream())


def clamp_enthalpy_ceiling_variable_cp(H, flags, cp_rho_field, H_liq, T_liquidus, T_vapor_cap,
                                       FLAG_GAS):
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(H), flags, cp_rho_field)
    _call("waam_clamp_enthalpy_ceiling_variable_cp", P(H), P(flags), P(cp_rho_field), H_liq,
          T_liquidus, T_vapor_cap, _n(H), _stream())


def compute_marangoni_force_variable(T, phi, f_l, Fx, Fy, Fz, flags, dgamma_lu_field, FLAG_SOLID,
                                     FLAG_GAS, nx, ny, nz):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, T, phi, f_l, Fx, Fy, Fz, flags, dgamma_lu_field)
    _call("waam_compute_marangoni_force_variable", P(T), P(phi), P(f_l), P(Fx), P(Fy), P(Fz),
          P(flags), P(dgamma_lu_field), nx, ny, nz, layout_of(T), _stream())


def remelt_hot_solid(T, H, f_l, phi, flags, cp_rho_field, T_solidus, T_liquidus, L_rho,
                     FLAG_SOLID, FLAG_FLUID):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_FLUID=FLAG_FLUID)
    _need(_n(T), H, f_l, phi, flags, cp_rho_field)
    _call("waam_remelt_hot_solid", P(T), P(H), P(f_l), P(phi), P(flags), P(cp_rho_field),
          T_solidus, T_liquidus, L_rho, _n(T), _stream())


def solidify_trailing_pool(T, H, f_l, phi, flags, ux, uy, uz, cp_rho_field, arc_i, arc_j, arc_k,
                           dir_x, dir_y, dir_z, lookback_cells, T_freeze, T_solidus, FLAG_SOLID,
                           FLAG_FLUID, FLAG_GAS):
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_FLUID=FLAG_FLUID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(T)
    _need(nx * ny * nz, H, f_l, phi, flags, ux, uy, uz, cp_rho_field)
    _call("waam_solidify_trailing_pool", P(T), P(H), P(f_l), P(phi), P(flags), P(ux), P(uy),
          P(uz), P(cp_rho_field), arc_i, arc_j, arc_k, dir_x, dir_y, dir_z, lookback_cells,
          T_freeze, T_solidus, nx, ny, nz, layout_of(T), _stream())


def extract_fl_yz_slice(f_l, out, i_slice, ny, nz):
    tensor_of(out).copy_(f_l[int(i_slice)] if isinstance(f_l, DeviceField)
                         else tensor_of(f_l)[int(i_slice)])


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("math", "torch")]


# ------------------------------------------------------------------ moving window
def shift_simulation_window_x(n_shift, f_a, f_b, T, H, f_l, phi, flags, T_max, T_prev, dT_dt,
                              time_above_800, time_above_1100, time_above_solidus, rho, ux, uy,
                              uz, Fx, Fy, Fz, cp_rho_field, alpha_lu_field, dgamma_lu_field,
                              tau_field, T_amb, cp_rho, rho0, nz_solid, FLAG_SOLID, FLAG_GAS,
                              FLAG_FLUID, nx, ny, nz):
    """Shift every field left by n_shift cells and reset the right strip (reference
    kernels.py:2620-2723)."""
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS, FLAG_FLUID=FLAG_FLUID)
    n = nx * ny * nz
    _need_pop(n, f_a, f_b)
    fields = (T, H, f_l, phi, flags, T_max, T_prev, dT_dt, time_above_800, time_above_1100,
              time_above_solidus, rho, ux, uy, uz, Fx, Fy, Fz, cp_rho_field, alpha_lu_field,
              dgamma_lu_field, tau_field)
    _need(n, *fields)
    lay = layout_of(T)
    if any(layout_of(x) != lay for x in (f_a, f_b) + fields):
        raise ValueError("shift_simulation_window_x: fields of different layouts")
    _call("waam_shift_simulation_window_x", int(n_shift), P(f_a), P(f_b),
          *(P(x) for x in fields), T_amb, cp_rho, rho0, int(nz_solid), nx, ny, nz, lay, _stream())


# ------------------------------------------------------------------ CTWD bead height
def bead_top_k(flags, nz_solid, out) -> None:
    """out[0] = largest k of a SOLID cell with k >= nz_solid, -1 when none (the reduction
    behind reference physics/electrical_stickout.py:39-50)."""
    nx, ny, nz = dims_of(flags)
    _call_ro("waam_bead_top_k", P(flags), int(nz_solid), P(out), nx, ny, nz, layout_of(flags),
             _stream())


# ------------------------------------------------------------------ full-tier surface forces
def apply_vapor_recoil(Fz, T, phi, flags, arc_i, arc_j, arc_k, sigma, recoil_pa, T_ref, dt, dx,
                       rho_ref, FLAG_SOLID, FLAG_GAS):
    """Reference kernels.py:995-1030."""
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(Fz)
    _need(nx * ny * nz, T, phi, flags)
    _call("waam_apply_vapor_recoil", P(Fz), P(T), P(phi), P(flags), arc_i, arc_j, arc_k, sigma,
          recoil_pa, T_ref, dt, dx, rho_ref, nx, ny, nz, layout_of(Fz), _stream())


def apply_vapor_recoil_clausius_clapeyron(Fz, T, phi, flags, arc_i, arc_j, arc_k, sigma, P_ref_Pa,
                                          T_boil_K, L_vapor_J_kg, R_spec_J_kgK, C_acc, dt, dx,
                                          rho_ref, FLAG_SOLID, FLAG_GAS):
    """Reference kernels.py:1052-1090."""
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(Fz)
    _need(nx * ny * nz, T, phi, flags)
    _call("waam_apply_vapor_recoil_clausius_clapeyron", P(Fz), P(T), P(phi), P(flags), arc_i,
          arc_j, arc_k, sigma, P_ref_Pa, T_boil_K, L_vapor_J_kg, R_spec_J_kgK, C_acc, dt, dx,
          rho_ref, nx, ny, nz, layout_of(Fz), _stream())


def apply_gas_shear_stress(Fx, Fy, phi, flags, arc_i, arc_j, arc_k, sigma, tau_peak_pa, dt, dx,
                           rho_ref, FLAG_SOLID, FLAG_GAS):
    """Reference kernels.py:1093-1139."""
    _flags_ok(FLAG_SOLID=FLAG_SOLID, FLAG_GAS=FLAG_GAS)
    nx, ny, nz = dims_of(Fx)
    _need(nx * ny * nz, Fy, phi, flags)
    _call("waam_apply_gas_shear_stress", P(Fx), P(Fy), P(phi), P(flags), arc_i, arc_j, arc_k,
          sigma, tau_peak_pa, dt, dx, rho_ref, nx, ny, nz, layout_of(Fx), _stream())


# ------------------------------------------------------------------ Lorentz force
def _scalar_out(buf):
    """Device address of a 0-d/1-element result buffer."""
    return P(buf)


def elec_build_sigma(sigma, f_l, flags, sigma_liquid, sigma_solid, FLAG_GAS):
    """Reference kernels.py:1236-1251."""
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(sigma), f_l, flags)
    _call("waam_elec_build_sigma", P(sigma), P(f_l), P(flags), sigma_liquid, sigma_solid,
          _n(sigma), _stream())


def elec_clear_source(source):
    _call("waam_elec_clear_source", P(source), _n(source), _stream())


def elec_inject_arc_source(source, arc_i, arc_j, arc_k, sigma_cells, current_A, dx, nz):
    """Reference kernels.py:1260-1286."""
    nx, ny, nz_ = dims_of(source)
    if nz != nz_:
        raise ValueError("nz does not match the source field")
    _call("waam_elec_inject_arc_source", P(source), arc_i, arc_j, arc_k, sigma_cells, current_A,
          dx, nx, ny, nz, layout_of(source), _stream())


def elec_normalize_source(source, target_current_A, dx, scratch=None):
    """Reference kernels.py:1289-1301."""
    sc = scratch if scratch is not None else _scratch_of(source)
    _call("waam_elec_normalize_source", P(source), target_current_A, dx, _n(source), P(sc),
          _stream())


def elec_init_ground(phi, flags, nz_solid, FLAG_SOLID):
    """Reference kernels.py:1304-1312 (phi = 0 everywhere)."""
    _flags_ok(FLAG_SOLID=FLAG_SOLID)
    _call("waam_elec_init_ground", P(phi), _n(phi), _stream())


def elec_jacobi_step(phi_in, phi_out, sigma, source, flags, dx, omega_j, FLAG_GAS, FLAG_SOLID,
                     nz_solid, nx, ny, nz):
    """Reference kernels.py:1315-1356."""
    _flags_ok(FLAG_GAS=FLAG_GAS, FLAG_SOLID=FLAG_SOLID)
    _need(nx * ny * nz, phi_in, phi_out, sigma, source, flags)
    _call("waam_elec_jacobi_step", P(phi_in), P(phi_out), P(sigma), P(source), P(flags), dx,
          omega_j, int(nz_solid), nx, ny, nz, layout_of(phi_in), _stream())


def elec_l1_diff_into(a, b, out, scratch=None):
    """out[0] = sum |a - b|, out[1] = sum |b| (reference kernels.py:1384-1396)."""
    sc = scratch if scratch is not None else _scratch_of(a)
    _call_ro("waam_elec_l1_diff", P(a), P(b), P(out), _n(a), P(sc), _stream())


def elec_compute_J(Jx, Jy, Jz, phi, sigma, flags, dx, FLAG_GAS, nx, ny, nz):
    """Reference kernels.py:1359-1381."""
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, Jx, Jy, Jz, phi, sigma, flags)
    _call("waam_elec_compute_J", P(Jx), P(Jy), P(Jz), P(phi), P(sigma), P(flags), dx, nx, ny, nz,
          layout_of(Jz), _stream())


def elec_bin_axial_current(Jz, bins, arc_i, arc_j, dx, bin_dr_cells, n_bins, nz):
    """Reference kernels.py:1399-1418; bins is (nz, n_bins)."""
    nx, ny, nz_ = dims_of(Jz)
    if nz != nz_ or _n(bins) < nz * n_bins:
        raise ValueError("bins must hold nz x n_bins values")
    _call("waam_elec_bin_axial_current", P(Jz), P(bins), arc_i, arc_j, dx, bin_dr_cells,
          int(n_bins), nx, ny, nz, layout_of(Jz), _stream())


def elec_prefix_bins(bins, n_bins, nz):
    """Reference kernels.py:1421-1432."""
    if _n(bins) < nz * n_bins:
        raise ValueError("bins must hold nz x n_bins values")
    _call("waam_elec_prefix_bins", P(bins), int(n_bins), int(nz), _stream())


def elec_B_axisymmetric(Bx, By, Bz, bins, arc_i, arc_j, dx, bin_dr_cells, mu0, n_bins, nz):
    """Reference kernels.py:1435-1477."""
    nx, ny, nz_ = dims_of(Bx)
    if nz != nz_ or _n(bins) < nz * n_bins:
        raise ValueError("bins must hold nz x n_bins values")
    _need(nx * ny * nz, By, Bz)
    _call("waam_elec_B_axisymmetric", P(Bx), P(By), P(Bz), P(bins), arc_i, arc_j, dx,
          bin_dr_cells, mu0, int(n_bins), nx, ny, nz, layout_of(Bx), _stream())


def apply_lorentz_JxB(Fx, Fy, Fz, Jx, Jy, Jz, Bx, By, Bz, f_l, flags, dt, dx, rho_ref, FLAG_GAS):
    """Reference kernels.py:1480-1513."""
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(_n(Fx), Fy, Fz, Jx, Jy, Jz, Bx, By, Bz, f_l, flags)
    _call("waam_apply_lorentz_JxB", P(Fx), P(Fy), P(Fz), P(Jx), P(Jy), P(Jz), P(Bx), P(By), P(Bz),
          P(f_l), P(flags), dt, dx, rho_ref, _n(Fx), _stream())


# ------------------------------------------------------------------ export-side fields
def compute_vorticity_magnitude(ux, uy, uz, flags, vort_out, dx, dt, FLAG_GAS, nx, ny, nz):
    """Reference kernels.py:3380-3419."""
    _flags_ok(FLAG_GAS=FLAG_GAS)
    _need(nx * ny * nz, ux, uy, uz, flags, vort_out)
    _call("waam_compute_vorticity_magnitude", P(ux), P(uy), P(uz), P(flags), P(vort_out), dx, dt,
          nx, ny, nz, layout_of(vort_out), _stream())


def max_norm3_into(fx, fy, fz, out) -> None:
    """out[0] = max |(fx, fy, fz)| (device reduction)."""
    _need(_n(fx), fy, fz)
    _call_ro("waam_max_norm3", P(fx), P(fy), P(fz), P(out), _n(fx), _stream())


def cavity_step(f_src, f_dst, rho, ux, uy, uz, flags, omega_s, omega_b, u_wall, nx, ny, nz):
    """One lid-driven cavity step (pull stream with bounce-back and a moving lid, MRT)."""
    _need(nx * ny * nz, rho, ux, uy, uz, flags)
    _need_pop(nx * ny * nz, f_src, f_dst)
    uwx, uwy, uwz = (float(v) for v in u_wall)
    _call("waam_cavity_step", P(f_src), P(f_dst), P(rho), P(ux), P(uy), P(uz), P(flags),
          float(omega_s), float(omega_b), uwx, uwy, uwz, nx, ny, nz, layout_of(rho), _stream())


def probe_gather_into(T, f_l, cells, out) -> None:
    """out[2p], out[2p + 1] = T, f_l at linear cell cells[p] (int64 device tensor, storage order
    of the fields; negative -> NaN)."""
    n = _n(T)
    _need(n, f_l)
    _call_ro("waam_probe_gather", P(T), P(f_l), P(cells), int(cells.numel()), n, P(out),
             _stream())
