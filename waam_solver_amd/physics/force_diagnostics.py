"""Per-force diagnostics: the max |F| (lattice units) of each force of the catalogue applied
alone to the current T / phi / f_l state (reference physics/force_diagnostics.py:1-150).

The live forces are kept in device copies and put back afterwards; each measurement is one
device max-norm reduction, and the nine maxima come back in one copy at the end (the reference
reads the three force volumes back to the host for every one).
"""

from __future__ import annotations

import math

import torch

from .. import kernels
from ..field import tensor_of
from . import weld_forces

KEYS = ("f_csf_max", "f_marangoni_max", "f_arc_max", "f_lorentz_max", "f_buoyancy_max",
        "f_gravity_max", "f_gas_shear_max", "f_recoil_max", "f_total_snap_max")


def _arc_ijk(twin) -> tuple[float, float, float]:
    """The last step's arc position; the arc height is resolved from the column when the step
    did not need it (a fused step without the arc)."""
    g = twin.grid
    last = getattr(twin, "_last_arc_ijk", None)
    if last is None:
        return (g.nx - 1) * 0.5, (g.ny - 1) * 0.5, float(twin.nz_solid)
    i, j, k = last
    if k is None:
        from ..solvers.coupled_step import resolve_arc_k

        k = resolve_arc_k(twin, g, i, j, None)
    return float(i), float(j), float(k)


def sample_force_diagnostics(twin) -> dict[str, float]:
    twin.sync()
    g = twin.grid
    n = g.nx, g.ny, g.nz
    arc_i, arc_j, arc_k = _arc_ijk(twin)
    saved = [tensor_of(f).clone() for f in (g.Fx, g.Fy, g.Fz)]
    res = torch.zeros(len(KEYS), dtype=torch.float32, device=saved[0].device)
    measured = set()

    def measure(key, apply_fn):
        kernels.clear_forces(g.Fx, g.Fy, g.Fz)
        apply_fn()
        kernels.max_norm3_into(g.Fx, g.Fy, g.Fz, res[KEYS.index(key)])
        measured.add(key)

    if twin.enable_csf_tension:
        measure("f_csf_max", lambda: kernels.compute_csf_tension(
            g.phi, g.flags, g.Fx, g.Fy, g.Fz, twin.gamma_lu, g.FLAG_SOLID, g.FLAG_GAS, *n,
            int(twin.enable_wetting), twin.theta_rad))
    if twin.use_material_tables:
        measure("f_marangoni_max", lambda: kernels.compute_marangoni_force_variable(
            g.T, g.phi, g.f_l, g.Fx, g.Fy, g.Fz, g.flags, g.dgamma_lu_field, g.FLAG_SOLID,
            g.FLAG_GAS, *n))
    else:
        measure("f_marangoni_max", lambda: kernels.compute_marangoni_force(
            g.T, g.phi, g.f_l, g.Fx, g.Fy, g.Fz, g.flags, twin.dgamma_dT_lu, g.dx, g.FLAG_SOLID,
            g.FLAG_GAS, *n))
    if twin.enable_gas_shear:
        measure("f_gas_shear_max", lambda: weld_forces.apply_gas_shear(twin, g, arc_i, arc_j,
                                                                       arc_k))
    p_arc = weld_forces.arc_pressure_peak_pa(twin)
    sig = weld_forces.arc_pressure_sigma_cells(twin)
    measure("f_arc_max", lambda: kernels.apply_arc_pressure(
        g.Fz, g.flags, g.phi, arc_i, arc_j, arc_k, sig, p_arc, g.dt, g.dx, twin.mat.rho,
        g.FLAG_SOLID, g.FLAG_GAS))
    if twin.enable_recoil:
        measure("f_recoil_max", lambda: weld_forces.apply_recoil(twin, g, arc_i, arc_j, arc_k))
    if twin.enable_hydrostatic_gravity:
        measure("f_gravity_max", lambda: kernels.add_hydrostatic_gravity(
            g.Fz, g.f_l, g.flags, 1.0, twin.g_lu, g.FLAG_SOLID, g.FLAG_GAS))
    measure("f_buoyancy_max", lambda: kernels.add_buoyancy(
        g.T, g.Fz, g.f_l, g.flags, twin.g_lu, twin.beta_T, twin.mat.T_liquidus, 1.0,
        g.FLAG_SOLID, g.FLAG_GAS))
    if twin.enable_lorentz:
        measure("f_lorentz_max", lambda: weld_forces.solve_lorentz(twin, g, arc_i, arc_j, arc_k))
    for f, s in zip((g.Fx, g.Fy, g.Fz), saved):
        tensor_of(f).copy_(s)
    kernels.max_norm3_into(g.Fx, g.Fy, g.Fz, res[KEYS.index("f_total_snap_max")])
    measured.add("f_total_snap_max")
    vals = res.cpu().tolist()
    return {k: (round(float(v), 6) if k in measured else 0.0) for k, v in zip(KEYS, vals)}


def diagnostics_have_nan(diag: dict) -> bool:
    return any(isinstance(v, float) and (math.isnan(v) or math.isinf(v)) for v in diag.values())
