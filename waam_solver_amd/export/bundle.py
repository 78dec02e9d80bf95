"""Research bundles and snapshot sequences.

On-disk format (the reference's export/bundle.py defines it; the names are what downstream tools
open):

    <bundle dir>/volume_<tag>.vti      ImageData volume of the selected tiers
                 surface_<tag>.vtp     phi = 0.5 isosurface (VOF runs only)
                 tracers_<tag>.vtp     porosity tracer points
                 telemetry_<tag>.json  get_telemetry(), indented JSON
                 meta_<tag>.json       sidecar (grid, units, physics switches, job)
                 probes.csv            probe rows, when probes are attached
    <sequence dir>/frame_NNNN/         one bundle per exported frame
                  sequence.pvd         ParaView collection of the frames' volumes

<tag> defaults to step_<step number, six digits>. In a sequence, a frame is exported after step
0, after every step whose 0-based index is a multiple of ``every_n``, and after the last step, up
to ``max_frames`` frames; frames 0, 5, 10, ... carry tracers. The collection lists each volume by
its path relative to the .pvd file (its bare file name when it lies outside that directory) and
stamps it with the simulation time of its step (the frame number when no times are given),
written with nine significant digits.
"""

from __future__ import annotations

import json
import logging
import pathlib
import xml.etree.ElementTree as ET
from dataclasses import dataclass
from typing import Callable, Sequence

from .meta import write_meta_json
from .vtk_io import TIER_MAP, export_surface, export_tracers, export_volume

log = logging.getLogger("waam_solver_amd")


@dataclass(frozen=True)
class BundleFiles:
    """The file names of one bundle directory."""
    root: pathlib.Path
    tag: str

    def named(self, stem: str, ext: str) -> pathlib.Path:
        return self.root / f"{stem}_{self.tag}.{ext}"

    @property
    def probes(self) -> pathlib.Path:
        return self.root / "probes.csv"


def _tier_codes(tiers: Sequence[int]) -> tuple[int, ...]:
    """Tier numbers 0-3 (core, material, forces, derived) as vtk_io's tier codes; other values
    pass through as codes."""
    return tuple(TIER_MAP[t] if t in TIER_MAP else t for t in tiers)


def export_research_bundle(twin, out_dir, tag: str | None = None,
                           tiers: Sequence[int] = (0, 1, 3), include_surface: bool = True,
                           include_tracers: bool = True, include_sidecar: bool = True,
                           crop_liquid: bool = False, job_path: str | None = None) -> dict[str, str]:
    """Write one snapshot of ``twin`` into ``out_dir``; returns {kind: path} for the files
    written (kinds: volume, surface, tracers, telemetry, meta, probes)."""
    files = BundleFiles(pathlib.Path(out_dir), tag or f"step_{twin._step_n:06d}")
    files.root.mkdir(parents=True, exist_ok=True)
    writers: list[tuple[str, bool, Callable[[], str | None]]] = [
        ("volume", True,
         lambda: export_volume(twin, str(files.named("volume", "vti")),
                               tiers=_tier_codes(tiers), crop_liquid=crop_liquid)),
        ("surface", include_surface and bool(twin.enable_vof),
         lambda: export_surface(twin, str(files.named("surface", "vtp")))),
        ("tracers", include_tracers,
         lambda: export_tracers(twin, str(files.named("tracers", "vtp")))),
        ("telemetry", True, lambda: _write_json(files.named("telemetry", "json"),
                                                twin.get_telemetry())),
        ("meta", include_sidecar, lambda: _write_meta(twin, files.named("meta", "json"),
                                                      job_path)),
        ("probes", _has_probes(twin), lambda: _write_probes(twin, files.probes)),
    ]
    written = {}
    for kind, wanted, write in writers:
        if not wanted:
            continue
        path = write()
        if path:
            written[kind] = str(path)
    log.info("research bundle: %d files in %s", len(written), files.root.resolve())
    return written


def _write_json(path: pathlib.Path, obj) -> pathlib.Path:
    path.write_text(json.dumps(obj, indent=2))
    return path


def _write_meta(twin, path: pathlib.Path, job_path: str | None) -> pathlib.Path:
    write_meta_json(twin, path, job_path=job_path)
    return path


def _has_probes(twin) -> bool:
    rec = getattr(twin, "probe_recorder", None)
    return rec is not None and bool(rec.probes)


def _write_probes(twin, path: pathlib.Path) -> pathlib.Path:
    twin.probe_recorder.write_csv(path)
    return path


def write_pvd(collection_path, vti_paths: list[str], times_s: list[float] | None = None) -> None:
    """ParaView collection (VTKFile type "Collection") of ``vti_paths``."""
    pvd = pathlib.Path(collection_path).resolve()
    root = ET.Element("VTKFile", {"type": "Collection", "version": "0.1"})
    coll = ET.SubElement(root, "Collection")
    for n, path in enumerate(vti_paths):
        stamp = times_s[n] if times_s is not None and n < len(times_s) else float(n)
        ET.SubElement(coll, "DataSet", {"timestep": f"{stamp:.9g}",
                                        "file": _relative_to(pathlib.Path(path), pvd.parent)})
    ET.indent(root, space="  ")
    pvd.write_text('<?xml version="1.0"?>\n' + ET.tostring(root, encoding="unicode"))


def _relative_to(path: pathlib.Path, base: pathlib.Path) -> str:
    absolute = (path if path.is_absolute() else pathlib.Path.cwd() / path).resolve()
    if absolute.is_relative_to(base):
        return absolute.relative_to(base).as_posix()
    return absolute.name


def _schedule(n_steps: int, every_n: int, max_frames: int) -> tuple[int, list[int]]:
    """(steps to take, 0-based step indices after which a frame is written). A frame is due
    after step 0, every multiple of ``every_n`` and the last step; once ``max_frames`` frames are
    written, the next due frame ends the run right after its step."""
    due = [s for s in range(n_steps) if s % every_n == 0 or s == n_steps - 1]
    if len(due) <= max_frames:
        return n_steps, due
    return due[max_frames] + 1, due[:max_frames]


def export_research_sequence(twin, out_dir, n_steps: int, every_n: int = 100,
                             torch_x_m: float = 0.01, torch_y_m: float | None = None,
                             is_welding: bool = True, max_frames: int = 200,
                             tiers: Sequence[int] = (0, 3), job_path: str | None = None,
                             after_frame: Callable | None = None) -> list[str]:
    """Step ``twin`` at a fixed torch position, writing a bundle per frame (schedule in the
    module docstring) and sequence.pvd; returns the frames' volume paths."""
    seq = pathlib.Path(out_dir)
    seq.mkdir(parents=True, exist_ok=True)
    g = twin.grid
    y = (g.ny // 2) * g.dx if torch_y_m is None else torch_y_m
    n_take, frame_steps = _schedule(n_steps, every_n, max_frames)
    frame_of = {s: n for n, s in enumerate(frame_steps)}
    volumes: list[str] = []
    stamps: list[float] = []
    for s in range(n_take):
        twin.step(torch_x_m, y, is_welding=is_welding)
        frame = frame_of.get(s)
        if frame is None:
            continue
        sub = seq / f"frame_{frame:04d}"
        paths = export_research_bundle(twin, sub, tag=f"step_{twin._step_n:06d}", tiers=tiers,
                                       include_tracers=frame % 5 == 0, job_path=job_path)
        if after_frame is not None:
            after_frame(frame, sub, paths)
        if "volume" in paths:
            volumes.append(paths["volume"])
            stamps.append(twin._step_n * g.dt)
    if volumes:
        write_pvd(seq / "sequence.pvd", volumes, times_s=stamps)
    return volumes
