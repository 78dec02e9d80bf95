"""Bead geometry telemetry: toe angle and reinforcement height (reference
physics/bead_geometry.py:13-47)."""

from __future__ import annotations

import math

import numpy as np


def estimate_toe_angle_deg(twin, g) -> float:
    """Contact-line slope from the phi step across the plate top in column (nx // 3, ny // 2):
    atan(1 / |dphi|) clamped to [10, 89] deg at the last cell with phi > 0.45 above the first
    (the contact angle when none). Only that column, cells nz_solid - 1 .. nz_solid + 7, is read
    back (one copy)."""
    i, j = g.nx // 3, g.ny // 2
    nz_ref = twin.nz_solid
    k0 = max(0, nz_ref - 1)
    k1 = min(g.nz - 1, nz_ref + 8)
    best = float(twin.contact_angle_deg)
    if k1 <= k0:
        return best
    phi = g.phi[i, j, k0:k1].cpu().numpy().astype(np.float32)
    for k in range(k0 + 1, k1):
        p = phi[k - k0]
        if p > 0.45:
            dphi = p - phi[k - k0 - 1]
            if abs(dphi) > 1e-4:
                slope = abs(dphi) * g.dx / max(g.dx, 1e-9)
                angle = math.degrees(math.atan(1.0 / max(slope, 1e-6)))
                best = min(89.0, max(10.0, angle))
    return best


def bead_reinforcement_height_mm(twin, g) -> float:
    """Solidified metal height above the substrate in mm."""
    from .electrical_stickout import measure_bead_height_m

    return measure_bead_height_m(twin, g) * 1000.0
