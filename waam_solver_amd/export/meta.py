"""JSON sidecar of a research bundle: grid, window, plate, material, physics switches, unit
conversions and the telemetry (reference export/meta.py:1-116; same keys)."""

from __future__ import annotations

import json
import pathlib
from typing import Any

WAAM_TWIN_VERSION = "2.0.0"


def _r(v, nd):
    return None if v is None else round(float(v), nd)


def build_meta_dict(twin, job_path: str | None = None) -> dict[str, Any]:
    g, mat = twin.grid, twin.mat
    telem = twin.get_telemetry()
    job = getattr(twin, "_job_config", None)
    ps, po = twin.plate_size_mm, twin.plate_origin_mm
    meta: dict[str, Any] = {
        "waam_twin_version": WAAM_TWIN_VERSION,
        "step": twin._step_n,
        "sim_time_ms": telem["sim_time_ms"],
        "grid": {"nx": g.nx, "ny": g.ny, "nz": g.nz, "dx_mm": round(g.dx * 1000.0, 6),
                 "dt_us": round(g.dt * 1e6, 4)},
        "origin_mm": [round(twin._window_offset_x_m * 1000.0, 4), 0.0, 0.0],
        "window_offset_x_mm": telem["window_offset_x_mm"],
        "preset": twin.preset_name,
        "plate_thickness_mm": _r(twin.plate_thickness_mm, 3),
        "plate_size_mm": None if not ps else [_r(ps[0], 3), _r(ps[1], 3)],
        "plate_origin_mm": None if not po else [_r(po[0], 3), _r(po[1], 3)],
        "plate_ij": list(twin.resolve_plate_ij()),
        "nz_solid": int(twin.nz_solid),
        "substrate_thickness_mm": round(float(twin.nz_solid) * g.dx * 1000.0, 3),
        "material": {"name": mat.name, "status": mat.status, "T_solidus_K": mat.T_solidus,
                     "T_liquidus_K": mat.T_liquidus, "rho_kgm3": mat.rho},
        "physics_flags": {k: getattr(twin, k) for k in (
            "enable_vof", "enable_csf_tension", "enable_recoil", "use_recoil_clausius_clapeyron",
            "enable_lorentz", "enable_gas_shear", "enable_droplet_impact_pressure",
            "enable_substrate_growth", "enable_moving_window", "enable_wetting",
            "enable_hydrostatic_gravity", "enable_bead_freeze", "enable_ctwd",
            "contact_angle_deg", "enable_heat_loss", "C_darcy")},
        "unit_conversions": {
            "velocity_lu_to_ms": "u_ms = u_lu * dx / dt",
            "force_lu_to_ms2": "a_ms2 = F_lu * dx / dt^2",
            "pressure_pa_gauge": "P = (rho_lu - 1) * cs^2 * rho_phys * (dx/dt)^2",
        },
        "telemetry": telem,
    }
    meta["physics_flags"]["heat_source"] = twin.heat_source_name
    meta["physics_flags"]["welding_current_A"] = twin.welding_current_A
    if job_path:
        meta["job_path"] = job_path
    if job:
        meta["process"] = job.get("process", {})
        meta["reference"] = job.get("reference", {})
    return meta


def write_meta_json(twin, path: str | pathlib.Path, job_path: str | None = None) -> None:
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        json.dump(build_meta_dict(twin, job_path), f, indent=2)
