"""Job files: torch path, plate/domain and process settings (reference job.py:9-404).

Covers what ``WAAMTwin.from_job`` and ``run_path`` need on the flow path (scalar or table materials):
physics tiers (thermal | flow | full), the process, deposition, wetting, arc, heat-loss,
interpass and calibration sections.
"""

from __future__ import annotations

import logging
import math
import pathlib
from typing import Any

log = logging.getLogger("waam_solver_amd")

PKG_ROOT = pathlib.Path(__file__).resolve().parent


def load_job_config(path) -> dict[str, Any]:
    import yaml

    p = pathlib.Path(path)
    if not p.exists():
        raise FileNotFoundError(f"Job file not found: {p}")
    with open(p) as f:
        job = yaml.safe_load(f) or {}
    job.setdefault("_base_dir", str(p.resolve().parent))
    return job


def parse_torch_path(job: dict[str, Any]) -> list[tuple[float, float, float]]:
    csv_ref = job.get("torch_path_csv")
    if csv_ref:
        from .torch_path import load_torch_path_csv

        p = pathlib.Path(csv_ref)
        if not p.is_absolute():
            # the job's directory, then the package directory (where the reference resolves
            # against its project root, reference job.py:18-25), then the working directory
            for base in (pathlib.Path(job.get("_base_dir", ".")), PKG_ROOT):
                if (base / p).exists():
                    p = base / p
                    break
        return load_torch_path_csv(p)
    return [(float(pt.get("x_mm", 0)) / 1000.0, float(pt.get("y_mm", 0)) / 1000.0,
             float(pt.get("z_mm", 0)) / 1000.0) for pt in (job.get("torch_path") or [])]


def resolve_plate_and_domain(job, preset_domain_mm, preset_dx_mm) -> dict[str, Any]:
    sim = job.get("simulation", {}) or {}
    plate = job.get("plate", {}) or {}
    domain = plate.get("domain_mm", sim.get("domain_mm"))
    domain_mm = tuple(float(v) for v in domain) if domain is not None else tuple(preset_domain_mm)
    dx = plate.get("dx_mm", sim.get("dx_mm"))
    thick = plate.get("thickness_mm", sim.get("plate_thickness_mm",
                                              sim.get("substrate_thickness_mm")))
    size = plate.get("size_mm", sim.get("plate_size_mm"))
    plate_size = None
    if size is not None:
        plate_size = (float(size[0]), float(size[1]))
    else:
        lx = plate.get("length_mm", sim.get("plate_length_mm"))
        ly = plate.get("width_mm", sim.get("plate_width_mm"))
        if lx is not None and ly is not None:
            plate_size = (float(lx), float(ly))
    origin = plate.get("origin_mm", sim.get("plate_origin_mm"))
    return {"domain_mm": domain_mm,
            "dx_mm": float(dx) if dx is not None else float(preset_dx_mm),
            "plate_thickness_mm": float(thick) if thick is not None else None,
            "plate_size_mm": plate_size,
            "plate_origin_mm": (float(origin[0]), float(origin[1])) if origin is not None else None}


def apply_physics_tier(twin, tier: str) -> None:
    key = str(tier or "flow").strip().lower()
    if key in ("base", "default", "standard_physics"):
        key = "flow"
    twin.physics_tier = key
    if key == "thermal":
        twin.enable_vof = twin.enable_csf_tension = twin.enable_wetting = False
        twin.enable_hydrostatic_gravity = twin.enable_bead_freeze = False
        twin.enable_lorentz = twin.enable_gas_shear = twin.enable_recoil = False
        twin.enable_droplet_impact_pressure = False
        twin.arc_pressure_model = "constant"
        return
    if key == "flow":
        twin.enable_vof = True
        twin.grid.ensure_vof_buffers()
        twin.enable_csf_tension = twin.enable_wetting = True
        twin.enable_hydrostatic_gravity = twin.enable_bead_freeze = True
        twin.enable_lorentz = twin.enable_gas_shear = twin.enable_recoil = False
        twin.enable_droplet_impact_pressure = True
        twin.arc_pressure_model = "constant"
        return
    if key != "full":
        log.warning(f"[job] unknown physics_tier '{tier}'; using 'full' (thermal | flow | full)")
    # full: flow tier + Lorentz force + gas shear + Lin-Eagar arc pressure; recoil stays as the
    # job sets it (reference job.py:104-123)
    twin.enable_vof = True
    twin.grid.ensure_vof_buffers()
    twin.enable_csf_tension = twin.enable_wetting = True
    twin.enable_hydrostatic_gravity = twin.enable_bead_freeze = True
    twin.enable_lorentz = True
    twin.grid.ensure_lorentz_fields()
    twin.enable_gas_shear = True
    twin.enable_droplet_impact_pressure = True
    twin.arc_pressure_model = "lin_eagar"
    twin.physics_tier = "full"


def apply_job_to_twin(twin, job: dict[str, Any]) -> None:
    process = job.get("process", {}) or {}
    wfs = process.get("wire_feed_m_min")
    if wfs is not None:
        drop_len = float(process.get("droplet_length_mm", 3.0)) / 1000.0
        if drop_len > 0:
            twin.droplet_freq = (float(wfs) / 60.0) / drop_len
        twin.wire_feed_m_s = float(wfs) / 60.0
    for key, attr, conv in (("wire_diameter_mm", "wire_diameter_m", lambda v: float(v) / 1000.0),
                            ("droplet_freq_hz", "droplet_freq", float),
                            ("pulse_frequency_hz", "pulse_frequency_hz", float),
                            ("droplet_transfer_mode", "droplet_transfer_mode", str),
                            ("transfer_mode", "droplet_transfer_mode", str),
                            ("droplet_size_jitter", "droplet_size_jitter", float),
                            ("impact_lead_angle_deg", "impact_lead_angle_deg", float),
                            ("current_A", "welding_current_A", float)):
        if key in process:
            setattr(twin, attr, conv(process[key]))
    if process.get("travel_speed_mm_s") is not None:
        twin.travel_speed_m_s = float(process["travel_speed_mm_s"]) / 1000.0
    adv = job.get("advanced_physics", {}) or {}
    for key, cast in (("gas_jet_velocity_m_s", float), ("gas_shear_coeff", float),
                      ("sigma_liquid_Sm", float), ("sigma_solid_Sm", float),
                      ("lorentz_jacobi_iters", int), ("T_boiling_K", float),
                      ("L_vapor_J_kg", float), ("R_spec_vapor_J_kgK", float),
                      ("recoil_accommodation", float)):
        if key in adv:
            setattr(twin, key, cast(adv[key]))
    sim = job.get("simulation", {}) or {}
    if "physics_tier" in sim:
        apply_physics_tier(twin, str(sim["physics_tier"]))
    if "strict_mode" in sim:
        twin.strict_mode = bool(sim["strict_mode"])
    for key in ("enable_recoil", "enable_csf_tension", "enable_lorentz", "enable_gas_shear",
                "enable_droplet_impact_pressure", "enable_vof", "enable_enthalpy_cap",
                "arc_surface_weighting", "enable_substrate_growth", "enable_wetting",
                "enable_hydrostatic_gravity", "enable_bead_freeze", "enable_ctwd"):
        if key in sim:
            setattr(twin, key, bool(sim[key]))
    if twin.enable_vof:
        twin.grid.ensure_vof_buffers()
    if sim.get("use_torch_z") or sim.get("enable_torch_z"):
        twin.use_torch_z = True
    wet = job.get("surface_wetting", {}) or {}
    if "contact_angle_deg" in wet:
        twin.contact_angle_deg = float(wet["contact_angle_deg"])
        twin.theta_rad = math.radians(twin.contact_angle_deg)
    dep = job.get("deposition", {}) or {}
    for key, attr in (("superheat_K", "deposition_superheat_K"),
                      ("footprint_sigma_scale", "deposition_footprint_sigma_scale"),
                      ("trailing_solidify_lookback_mm", "trailing_solidify_lookback_mm"),
                      ("trailing_solidify_temp_margin_K", "trailing_solidify_temp_margin_K")):
        if key in dep:
            setattr(twin, attr, float(dep[key]))
    arc = job.get("arc_physics", {}) or {}
    if "sigma_mm" in arc:
        twin.arc_sigma_m = float(arc["sigma_mm"]) / 1000.0
        twin.sigma_cells = twin.arc_sigma_m / twin.grid.dx
    if "arc_sigma_mm" in process:
        twin.arc_sigma_m = float(process["arc_sigma_mm"]) / 1000.0
        twin.sigma_cells = twin.arc_sigma_m / twin.grid.dx
    if "penetration_mm" in arc:
        twin.arc_penetration_m = float(arc["penetration_mm"]) / 1000.0
    if "T_vapor_cap_K" in arc:
        twin.T_vapor_cap_K = float(arc["T_vapor_cap_K"])
    if "surface_weighting" in arc:
        twin.arc_surface_weighting = bool(arc["surface_weighting"])
    if "pressure_model" in arc:
        twin.arc_pressure_model = str(arc["pressure_model"]).strip().lower()
    if "pressure_pa" in arc:
        twin.arc_pressure = float(arc["pressure_pa"])
    if arc.get("pressure_sigma_mm") is not None:
        twin.pressure_sigma_m = float(arc["pressure_sigma_mm"]) / 1000.0
    if job.get("heat_source"):
        from .physics.arc import create_heat_source

        gold = job.get("goldak", {}) or {}
        name = str(job["heat_source"])
        key = name.lower().replace("-", "").replace("_", "")
        if key in ("goldak", "goldak3d", "doubleellipsoid"):
            dx_mm = twin.grid.dx * 1000.0
            cells = lambda k: max(0.5, float(gold[k]) / dx_mm) if gold.get(k) is not None else 0.0  # noqa
            src = create_heat_source(name, ff=gold.get("ff", 0.6), fr=gold.get("fr", 1.4))
            src.a_front, src.a_rear = cells("a_front_mm"), cells("a_rear_mm")
            src.b, src.c = cells("b_mm"), cells("c_mm")
            twin.arc_source = src
        else:
            twin.arc_source = create_heat_source(name)
        twin.heat_source_name = name
    interpass = job.get("interpass", {}) or {}
    if interpass:
        twin._interpass_cooling_steps = int(interpass.get("cooling_steps", 0))
    heat = job.get("heat_loss", {}) or {}
    if heat:
        twin.enable_convection = bool(heat.get("convection", False))
        twin.enable_radiation = bool(heat.get("radiation", False))
        twin.enable_heat_loss = twin.enable_convection or twin.enable_radiation
        if "h_conv" in heat:
            twin.h_conv = float(heat["h_conv"])
        if "eps_rad" in heat:
            twin.eps_rad = float(heat["eps_rad"])
    cal_path = job.get("calibration")
    if cal_path:
        from .calibration import apply_calibration, load_calibration

        apply_calibration(twin, load_calibration(cal_path))
