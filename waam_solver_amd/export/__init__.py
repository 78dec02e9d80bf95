"""On-disk formats of a run: VTK ImageData volumes (.vti) with the reference's array names,
tracer point clouds (.vtp), JSON telemetry and sidecars, probe CSV, research bundles and .pvd
collections (reference export/__init__.py)."""

from .bundle import export_research_bundle, export_research_sequence, write_pvd
from .meta import build_meta_dict, write_meta_json
from .probes import ProbeRecorder, ProbeSpec
from .vtk_io import (TIER_CORE, TIER_DERIVED, TIER_FORCES, TIER_MATERIAL, export_legacy_minimal,
                     export_surface, export_tracers, export_volume)

__all__ = ["ProbeRecorder", "ProbeSpec", "TIER_CORE", "TIER_MATERIAL", "TIER_FORCES",
           "TIER_DERIVED", "build_meta_dict", "write_meta_json", "export_volume",
           "export_surface", "export_tracers", "export_legacy_minimal", "export_research_bundle",
           "export_research_sequence", "write_pvd"]
