"""Research bundles: volume + tracers + telemetry + sidecar + probes per snapshot, scheduled
sequences and the ParaView .pvd collection (reference export/bundle.py:1-170)."""

from __future__ import annotations

import json
import logging
import pathlib
from typing import Callable, Sequence

from .meta import write_meta_json
from .vtk_io import TIER_MAP, export_surface, export_tracers, export_volume

log = logging.getLogger("waam_solver_amd")


def _parse_tiers(tiers: Sequence[int]) -> tuple[int, ...]:
    return tuple(TIER_MAP.get(t, t) for t in tiers)


def export_research_bundle(twin, out_dir, tag: str | None = None,
                           tiers: Sequence[int] = (0, 1, 3), include_surface: bool = True,
                           include_tracers: bool = True, include_sidecar: bool = True,
                           crop_liquid: bool = False, job_path: str | None = None) -> dict[str, str]:
    """{kind: path} of the files written to ``out_dir``: volume_<tag>.vti, surface_<tag>.vtp
    (with VOF), tracers_<tag>.vtp, telemetry_<tag>.json, meta_<tag>.json, probes.csv."""
    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    step_tag = tag or f"step_{twin._step_n:06d}"
    paths: dict[str, str] = {}
    p = export_volume(twin, str(out_dir / f"volume_{step_tag}.vti"), tiers=_parse_tiers(tiers),
                      crop_liquid=crop_liquid)
    if p:
        paths["volume"] = p
    if include_surface and twin.enable_vof:
        p = export_surface(twin, str(out_dir / f"surface_{step_tag}.vtp"))
        if p:
            paths["surface"] = p
    if include_tracers:
        p = export_tracers(twin, str(out_dir / f"tracers_{step_tag}.vtp"))
        if p:
            paths["tracers"] = p
    telem_path = out_dir / f"telemetry_{step_tag}.json"
    with open(telem_path, "w") as f:
        json.dump(twin.get_telemetry(), f, indent=2)
    paths["telemetry"] = str(telem_path)
    if include_sidecar:
        meta_path = out_dir / f"meta_{step_tag}.json"
        write_meta_json(twin, meta_path, job_path=job_path)
        paths["meta"] = str(meta_path)
    rec = twin.probe_recorder
    if rec is not None and rec.probes:
        probe_path = out_dir / "probes.csv"
        rec.write_csv(probe_path)
        paths["probes"] = str(probe_path)
    log.info(f"[export] research bundle -> {out_dir.resolve()} ({len(paths)} files)")
    return paths


def write_pvd(collection_path, vti_paths: list[str], times_s: list[float] | None = None) -> None:
    """A ParaView collection of the snapshots, paths relative to the .pvd's directory and
    timestep = the simulation time in s (the frame index without ``times_s``)."""
    collection_path = pathlib.Path(collection_path).resolve()
    base = collection_path.parent
    lines = ['<?xml version="1.0"?>', '<VTKFile type="Collection" version="0.1">',
             "  <Collection>"]
    for n, vti in enumerate(vti_paths):
        v = pathlib.Path(vti)
        v = (v if v.is_absolute() else pathlib.Path.cwd() / v).resolve()
        try:
            rel = v.relative_to(base)
        except ValueError:
            rel = pathlib.Path(v.name)
        t = times_s[n] if times_s is not None and n < len(times_s) else float(n)
        lines.append(f'    <DataSet timestep="{t:.9g}" file="{rel.as_posix()}"/>')
    lines += ["  </Collection>", "</VTKFile>"]
    collection_path.write_text("\n".join(lines))


def export_research_sequence(twin, out_dir, n_steps: int, every_n: int = 100,
                             torch_x_m: float = 0.01, torch_y_m: float | None = None,
                             is_welding: bool = True, max_frames: int = 200,
                             tiers: Sequence[int] = (0, 3), job_path: str | None = None,
                             after_frame: Callable | None = None) -> list[str]:
    """Step the twin and write a bundle every ``every_n`` steps (and after the last), tracers
    on frame 0 and every fifth frame, then sequence.pvd (reference export/bundle.py:117-170)."""
    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    g = twin.grid
    if torch_y_m is None:
        torch_y_m = (g.ny // 2) * g.dx
    vti, times = [], []
    frame = 0
    for s in range(n_steps):
        twin.step(torch_x_m, torch_y_m, is_welding=is_welding)
        if s % every_n != 0 and s != n_steps - 1:
            continue
        if frame >= max_frames:
            break
        sub = out_dir / f"frame_{frame:04d}"
        paths = export_research_bundle(twin, sub, tag=f"step_{twin._step_n:06d}", tiers=tiers,
                                       include_tracers=(frame % 5 == 0), job_path=job_path)
        if after_frame is not None:
            after_frame(frame, sub, paths)
        if "volume" in paths:
            vti.append(paths["volume"])
            times.append(twin._step_n * g.dt)
        frame += 1
    if vti:
        write_pvd(out_dir / "sequence.pvd", vti, times_s=times)
    return vti
