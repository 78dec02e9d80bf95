"""Scalar calibration overlays (reference calibration.py:1-47).

A calibration profile rescales four process parameters of a twin: the arc efficiency eta, the
convective coefficient h_conv (heat_loss_factor), the Marangoni coefficient (marangoni_scale;
read by the temperature-dependent material path) and the arc width in cells
(arc_sigma_scale). These are the parameters the differentiation form fits.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path


@dataclass
class CalibrationProfile:
    arc_efficiency: float = 1.0
    heat_loss_factor: float = 1.0
    marangoni_scale: float = 1.0
    arc_sigma_scale: float = 1.0


def _resolve(path) -> Path:
    p = Path(path)
    if p.is_absolute() or p.exists():
        return p
    pkg = Path(__file__).resolve().parent / p
    return pkg if pkg.exists() else p


def load_calibration(path) -> CalibrationProfile:
    """Profile from a YAML file with any of the four keys (missing keys: 1.0); None or '' gives
    the identity profile (reference calibration.py:18-35)."""
    if not path:
        return CalibrationProfile()
    import yaml

    with open(_resolve(path)) as fh:
        data = yaml.safe_load(fh) or {}
    return CalibrationProfile(
        arc_efficiency=float(data.get("arc_efficiency", 1.0)),
        heat_loss_factor=float(data.get("heat_loss_factor", 1.0)),
        marangoni_scale=float(data.get("marangoni_scale", 1.0)),
        arc_sigma_scale=float(data.get("arc_sigma_scale", 1.0)),
    )


def apply_calibration(twin, cal: CalibrationProfile) -> None:
    """Overlay the profile on a twin (reference calibration.py:38-47): eta is replaced, h_conv
    and sigma_cells are multiplied, marangoni_scale is set; factors equal to 1 change nothing."""
    if cal.arc_efficiency != 1.0:
        twin.eta = cal.arc_efficiency
    if cal.heat_loss_factor != 1.0:
        twin.h_conv *= cal.heat_loss_factor
    if cal.marangoni_scale != 1.0:
        twin.marangoni_scale = cal.marangoni_scale
    if cal.arc_sigma_scale != 1.0:
        twin.sigma_cells *= cal.arc_sigma_scale


__all__ = ["CalibrationProfile", "load_calibration", "apply_calibration"]
