"""Pool width/depth from a full f_l readback (reference benchmark.py:8-34)."""

from __future__ import annotations

import numpy as np


def measure_pool_mm(twin) -> tuple[float, float, int]:
    g = twin.grid
    liquid = g.f_l.to_numpy() > 0.5
    n_liq = int(liquid.sum())
    if not liquid.any():
        return 0.0, 0.0, 0
    i_c = int(round(float(np.nonzero(liquid)[0].mean())))
    sect = liquid[i_c]
    if not sect.any():
        return 0.0, 0.0, n_liq
    y_idx = np.where(sect.any(axis=1))[0]
    z_idx = np.where(sect.any(axis=0))[0]
    W = (y_idx[-1] - y_idx[0] + 1) * g.dx * 1000.0
    D = max(0, int(getattr(twin, "nz_solid", 0)) - z_idx[0]) * g.dx * 1000.0
    return float(W), float(D), n_liq
