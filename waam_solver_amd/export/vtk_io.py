"""VTK volume and tracer export (reference export/vtk_io.py:1-267).

The array set, names and units follow the reference's ``export_volume`` tiers:
  TIER_CORE     temperature (K, C), liquid fraction, VOF phi, cell flags, T_max, T_prev, dT/dt,
                cooling rate (-dT/dt), enthalpy, the three times above thresholds, velocity
                (m/s) and speed, lattice density and gauge pressure;
  TIER_MATERIAL the per-cell rho cp, alpha_lu, dgamma/dT_lu and SRT tau (table materials);
  TIER_FORCES   the body force (m/s^2) and, with the Lorentz force, J and B;
  TIER_DERIVED  curvature, vorticity magnitude (1/s) and the mushy-zone mask.
``export_surface`` writes the phi = 0.5 free surface as triangles. Fields are read back in one
batch: every array is copied into one pinned host buffer with
asynchronous copies and a single synchronisation (the twin's (nz, ny, nx) storage already is
VTK's cell order, so no transpose is needed there). Files are written by ``vtkxml`` without
pyvista.
"""

from __future__ import annotations

import logging
import os
import pathlib
from typing import Sequence

import numpy as np
import torch

from .. import kernels
from ..field import LAYOUT_ZYX, tensor_of
from .isosurface import cell_to_point, marching_tetrahedra
from .vtkxml import write_image_data, write_poly_data, write_poly_points

log = logging.getLogger("waam_solver_amd")

TIER_CORE, TIER_MATERIAL, TIER_FORCES, TIER_DERIVED = 0, 1, 2, 3
TIER_MAP = {0: TIER_CORE, 1: TIER_MATERIAL, 2: TIER_FORCES, 3: TIER_DERIVED}


def _headless() -> bool:
    return os.environ.get("WAAM_HEADLESS") == "1"


def vtk_imagedata_path(path: str) -> str:
    p = pathlib.Path(path)
    if p.suffix.lower() in (".vts", ""):
        return str(p.with_suffix(".vti"))
    return str(path)


def readback_vtk_order(fields: dict) -> dict[str, np.ndarray]:
    """{name: flat numpy array in VTK cell order (x fastest)} for grid DeviceFields, through one
    pinned host buffer and one synchronisation."""
    views = {}
    for name, f in fields.items():
        t = tensor_of(f)
        if getattr(f, "layout", None) != LAYOUT_ZYX and t.dim() >= 3:
            t = t.permute(*range(t.dim() - 3), t.dim() - 1, t.dim() - 2, t.dim() - 3)
        views[name] = t.contiguous().reshape(-1)
    out = {}
    if not views:
        return out
    pad = lambda nb: (nb + 255) // 256 * 256  # noqa: E731  every slot starts 256-B aligned
    total = sum(pad(v.numel() * v.element_size()) for v in views.values())
    host = torch.empty(total, dtype=torch.uint8, pin_memory=torch.cuda.is_available())
    off = 0
    slots = []
    for name, v in views.items():
        nb = v.numel() * v.element_size()
        dst = host[off:off + nb].view(v.dtype)
        dst.copy_(v, non_blocking=True)
        slots.append((name, off, nb, v.dtype))
        off += pad(nb)
    if torch.cuda.is_available():
        torch.cuda.current_stream().synchronize()
    raw = host.numpy()
    for name, o, nb, dt in slots:
        out[name] = raw[o:o + nb].view(torch.empty(0, dtype=dt).numpy().dtype).copy()
    return out


def _prepare_derived(twin, tiers) -> None:
    if TIER_DERIVED not in tiers:
        return
    g = twin.grid
    g.ensure_export_buffers()
    kernels.compute_curvature_field(g.phi, g.flags, g.kappa_field, g.FLAG_GAS, g.nx, g.ny, g.nz)
    kernels.compute_vorticity_magnitude(g.ux, g.uy, g.uz, g.flags, g.vorticity_mag, g.dx, g.dt,
                                        g.FLAG_GAS, g.nx, g.ny, g.nz)


def _force_fields(twin):
    """The step's body forces: the fused engine keeps them in Fx/Fy/Fz, the other engine in the
    snapshot fields."""
    g = twin.grid
    if twin.engine == "fused" and not (twin.use_material_tables or twin.enable_lorentz):
        return g.Fx, g.Fy, g.Fz
    return g.Fx_snap, g.Fy_snap, g.Fz_snap


def export_volume(twin, path: str, tiers: Sequence[int] = (TIER_CORE, TIER_DERIVED),
                  crop_liquid: bool = False) -> str | None:
    """Write the volume as VTK ImageData (.vti), cell data by tier (reference
    export/vtk_io.py:85-177)."""
    if _headless():
        return None
    if crop_liquid:
        raise NotImplementedError("crop_liquid (an unstructured threshold of the volume) is not "
                                  "built; threshold Liquid_Fraction in the viewer instead")
    twin.sync()
    g = twin.grid
    tiers = tuple(TIER_MAP.get(t, t) for t in tiers)
    _prepare_derived(twin, tiers)
    want = {"T": g.T, "f_l": g.f_l}
    if TIER_CORE in tiers:
        want.update(phi=g.phi, flags=g.flags, T_max=g.T_max, T_prev=g.T_prev, dT_dt=g.dT_dt,
                    H=g.H, t800=g.time_above_800_s, t1100=g.time_above_1100_s,
                    tsol=g.time_above_solidus_s, ux=g.ux, uy=g.uy, uz=g.uz, rho=g.rho)
    if TIER_MATERIAL in tiers and twin.use_material_tables:
        want.update(cp=g.cp_rho_field, alpha=g.alpha_lu_field, dgamma=g.dgamma_lu_field,
                    tau=g.tau_field)
    if TIER_FORCES in tiers:
        fx, fy, fz = _force_fields(twin)
        want.update(Fx=fx, Fy=fy, Fz=fz)
        if twin.enable_lorentz:
            g.ensure_lorentz_fields()
            want.update(Jx=g.Jx, Jy=g.Jy, Jz=g.Jz, Bx=g.Bx, By=g.By, Bz=g.Bz)
    if TIER_DERIVED in tiers:
        want.update(kappa=g.kappa_field, vort=g.vorticity_mag)
    a = readback_vtk_order(want)
    to_ms = lambda x: x * g.dx / g.dt  # noqa: E731  lattice velocity -> m/s
    cd: dict[str, np.ndarray] = {}
    if TIER_CORE in tiers:
        T = a["T"]
        ux, uy, uz = to_ms(a["ux"]), to_ms(a["uy"]), to_ms(a["uz"])
        cd["Temperature_K"] = T
        cd["Temperature_C"] = T - 273.15
        cd["Liquid_Fraction"] = a["f_l"]
        cd["VOF_phi"] = a["phi"]
        cd["Cell_Flags"] = a["flags"].astype(np.int32)
        cd["T_max_K"] = a["T_max"]
        cd["T_prev_K"] = a["T_prev"]
        cd["dTdt_Ks"] = a["dT_dt"]
        cd["Cooling_Rate_Ks"] = -a["dT_dt"]
        cd["Enthalpy_Jm3"] = a["H"]
        cd["Time_above_800C_s"] = a["t800"]
        cd["Time_above_1100C_s"] = a["t1100"]
        cd["Time_above_solidus_s"] = a["tsol"]
        cd["Velocity_X_ms"], cd["Velocity_Y_ms"], cd["Velocity_Z_ms"] = ux, uy, uz
        cd["Speed_ms"] = np.sqrt(ux ** 2 + uy ** 2 + uz ** 2)
        cd["Density_lu"] = a["rho"]
        # gauge pressure p = cs^2 (rho_lu - 1) rho_phys (dx/dt)^2
        cd["Pressure_Pa_gauge"] = (a["rho"] - 1.0) * (1.0 / 3.0) * g.mat.rho * (g.dx / g.dt) ** 2
    if "cp" in a:
        cd["RhoCp_Jm3K"], cd["Alpha_lu"] = a["cp"], a["alpha"]
        cd["DgammaDT_lu"], cd["Tau_SRT"] = a["dgamma"], a["tau"]
    if TIER_FORCES in tiers:
        fx, fy, fz = (to_ms(a[k]) / g.dt for k in ("Fx", "Fy", "Fz"))
        cd["BodyForce_X_ms2"], cd["BodyForce_Y_ms2"], cd["BodyForce_Z_ms2"] = fx, fy, fz
        cd["BodyForce_mag_ms2"] = np.sqrt(fx ** 2 + fy ** 2 + fz ** 2)
        if "Jx" in a:
            cd["CurrentDensity_X_Am2"], cd["CurrentDensity_Y_Am2"] = a["Jx"], a["Jy"]
            cd["CurrentDensity_Z_Am2"] = a["Jz"]
            cd["MagneticField_X_T"], cd["MagneticField_Y_T"] = a["Bx"], a["By"]
            cd["MagneticField_Z_T"] = a["Bz"]
    if TIER_DERIVED in tiers:
        fl = a["f_l"]
        cd["Curvature_kappa"] = a["kappa"]
        cd["Vorticity_mag_1s"] = a["vort"]
        cd["Mushy_Zone"] = ((fl > 0.05) & (fl < 0.95)).astype(np.float32)
    out = vtk_imagedata_path(path)
    write_image_data(out, (g.nx, g.ny, g.nz), g.dx * 1000.0,
                     (twin._window_offset_x_m * 1000.0, 0.0, 0.0), cd)
    log.info(f"[export] volume VTK -> {out} ({len(cd)} arrays)")
    return out


def export_surface(twin, path: str, include_kappa: bool = True) -> str | None:
    """The phi = 0.5 free surface (the bead profile) as a triangle PolyData with phi,
    Liquid_Fraction, Temperature_K and Curvature_kappa interpolated onto it; the f_l = 0.5 melt
    front when phi has no such level (reference export/vtk_io.py:180-229). Cell data is averaged
    to the grid points and contoured by marching tetrahedra (``isosurface``); coordinates are
    in mm with the window offset, as the volume's."""
    if _headless():
        return None
    twin.sync()
    g = twin.grid
    want = {"phi": g.phi, "fl": g.f_l, "T": g.T}
    if include_kappa:
        g.ensure_export_buffers()
        kernels.compute_curvature_field(g.phi, g.flags, g.kappa_field, g.FLAG_GAS, g.nx, g.ny,
                                        g.nz)
        want["kappa"] = g.kappa_field
    a = readback_vtk_order(want)
    shape_zyx = (g.nz, g.ny, g.nx)
    pts = {k: cell_to_point(v.reshape(shape_zyx).transpose(2, 1, 0)) for k, v in a.items()}
    names = {"phi": "phi", "fl": "Liquid_Fraction", "T": "Temperature_K",
             "kappa": "Curvature_kappa"}
    arrays = {names[k]: v for k, v in pts.items()}
    verts = tris = data = None
    for scalar in ("phi", "fl"):
        verts, tris, data = marching_tetrahedra(pts[scalar], 0.5, arrays)
        if len(tris):
            break
    if tris is None or len(tris) == 0:
        log.info("[export] no phi/f_l = 0.5 surface found; skipping surface VTK")
        return None
    xyz = verts * (g.dx * 1000.0)
    xyz[:, 0] += twin._window_offset_x_m * 1000.0
    write_poly_data(path, xyz, {k: v.astype(np.float32) for k, v in data.items()},
                    triangles=tris)
    log.info(f"[export] surface VTK -> {path} ({len(tris)} triangles)")
    return str(path)


def export_tracers(twin, path: str) -> str | None:
    """Active and trapped porosity tracers as a PolyData point cloud in mm, with the window
    offset added back (reference export/vtk_io.py:232-260)."""
    if _headless():
        return None
    g = twin.grid
    pos = g.porosity_pos.to_numpy()
    active = g.porosity_active.to_numpy()
    mask = active > 0
    if not mask.any():
        log.info("[export] no active tracers; skipping tracer VTK")
        return None
    pts = pos[mask].copy()
    pts[:, 0] += twin._window_offset_x_m
    pts *= 1000.0
    st = active[mask].astype(np.int32)
    write_poly_points(path, pts, {"state": st, "trapped": (st == 2).astype(np.int32)})
    log.info(f"[export] tracers VTK -> {path} ({pts.shape[0]} points)")
    return str(path)


def export_legacy_minimal(twin, path: str) -> str | None:
    """The core tier only (reference export/vtk_io.py:263-265)."""
    return export_volume(twin, path, tiers=(TIER_CORE,), crop_liquid=False)
