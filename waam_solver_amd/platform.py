"""Presets, grid sizing and device query (reference platform.py).

The reference picks a Taichi backend (CUDA, then Vulkan, then CPU); here the backend is fixed —
ROCm/HIP on the current GPU — and the memory budget of a preset is checked against the device's
HBM as reported by torch.
"""

from __future__ import annotations

import os
import pathlib
from dataclasses import dataclass

_PRESETS = pathlib.Path(__file__).resolve().parent / "config" / "presets.yaml"
# bytes per cell counted by the reference's budget (platform.py:15)
_BYTES_PER_CELL = 2 * 19 * 4 + 7 * 4 + 6 * 4 + 4


@dataclass
class PresetConfig:
    name: str
    vram_budget_mb: int
    domain_mm: tuple[float, float, float]
    target_dx_mm: float
    max_tracers: int
    use_srt: bool


def load_presets() -> dict[str, PresetConfig]:
    import yaml

    with open(_PRESETS) as f:
        raw = yaml.safe_load(f)
    out = {}
    for name, cfg in raw.items():
        d = cfg.get("domain_mm", [80, 40, 25])
        out[name] = PresetConfig(name, int(cfg.get("vram_budget_mb", 2048)),
                                 (float(d[0]), float(d[1]), float(d[2])),
                                 float(cfg.get("target_dx_mm", 0.3)),
                                 int(cfg.get("max_tracers", 20000)), bool(cfg.get("use_srt", True)))
    return out


def device_memory_mb() -> int | None:
    override = os.environ.get("WAAM_VRAM_MB")
    if override:
        return int(override)
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.get_device_properties(0).total_memory // (1024 ** 2))
    except Exception:
        pass
    return None


def _auto_tier() -> str:
    v = device_memory_mb()
    if v is None or v < 1024:
        return "minimal"
    if v < 6144:
        return "standard"
    if v < 12288:
        return "high"
    return "ultra"


def resolve_preset(name: str | None = None) -> PresetConfig:
    key = (name or os.environ.get("WAAM_PRESET", "standard")).lower()
    if key == "auto":
        key = _auto_tier()
    presets = load_presets()
    if key not in presets:
        raise KeyError(f"Unknown preset '{key}'. Available: {list(presets)}")
    return presets[key]


def auto_tracer_count(vram_mb: int, cfg: PresetConfig) -> int:
    return int(cfg.max_tracers)


def estimate_grid_vram_mb(nx: int, ny: int, nz: int, max_tracers: int) -> float:
    return (nx * ny * nz * _BYTES_PER_CELL + max_tracers * 16) / (1024 ** 2)


def auto_grid(domain_mm, target_dx_mm: float, vram_budget_mb: int, max_tracers: int = 20000):
    """(nx, ny, nz, dx): coarsen dx by 1.15 while the grid exceeds 85 % of the budget."""
    dx_m = target_dx_mm / 1000.0
    lx, ly, lz = (d / 1000.0 for d in domain_mm)
    dims = lambda: (max(8, int(lx / dx_m)), max(8, int(ly / dx_m)), max(8, int(lz / dx_m)))  # noqa
    nx, ny, nz = dims()
    budget = float(vram_budget_mb) * 0.85
    while estimate_grid_vram_mb(nx, ny, nz, max_tracers) > budget and dx_m < 0.002:
        dx_m *= 1.15
        nx, ny, nz = dims()
    est = estimate_grid_vram_mb(nx, ny, nz, max_tracers)
    if est > budget:
        raise MemoryError(f"Grid {nx}x{ny}x{nz} needs ~{est:.0f} MB but budget is {vram_budget_mb} MB")
    return nx, ny, nz, dx_m


def check_vram_budget(nx, ny, nz, max_tracers, budget_mb) -> None:
    est = estimate_grid_vram_mb(nx, ny, nz, max_tracers)
    if est > budget_mb:
        raise MemoryError(f"Estimated memory {est:.1f} MB exceeds budget {budget_mb} MB")
    dev = device_memory_mb()
    if dev is not None and est > dev:
        raise MemoryError(f"Estimated memory {est:.1f} MB exceeds the device's {dev} MB")
