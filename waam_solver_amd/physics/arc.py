"""Arc heat sources (reference physics/arc.py:27-213)."""

from __future__ import annotations

import logging
from dataclasses import dataclass

from .. import kernels

log = logging.getLogger("waam_solver_amd")


def _arc_weight_args(twin) -> tuple[float, int]:
    return twin.arc_penetration_m / twin.grid.dx, 1 if twin.arc_surface_weighting else 0


def normalize_goldak_fractions(ff: float, fr: float) -> tuple[float, float]:
    """Enforce the Goldak convention f_f + f_r = 2."""
    s = float(ff) + float(fr)
    if s <= 1e-12:
        log.warning("[Goldak] ff+fr<=0; using ff=0.6, fr=1.4")
        return 0.6, 1.4
    if abs(s - 2.0) <= 0.05:
        return float(ff), float(fr)
    if abs(s - 1.0) <= 0.05:
        return 2.0 * float(ff), 2.0 * float(fr)
    return 2.0 * float(ff) / s, 2.0 * float(fr) / s


@dataclass
class Gaussian2D:
    """Surface Gaussian: weights exp(-(di^2 + dj^2 + dk^2/4) / 2 sigma^2)."""

    def inject(self, twin, g, arc_i, arc_j, arc_k) -> None:
        pen, en = _arc_weight_args(twin)
        kernels.inject_arc_heat(g.H, g.flags, g.phi, g.f_l, g.arc_norm_buf, arc_i, arc_j, arc_k,
                                twin.Q_w, twin.sigma_cells, g.dt, g.dx ** 3, twin.eta, pen, en,
                                g.FLAG_SOLID, g.FLAG_GAS, scratch=g.scratch)


@dataclass
class Goldak3D:
    """Goldak double ellipsoid; semi-axes in cells (<= 0: resolved from the arc sigma)."""

    ff: float = 0.6
    fr: float = 1.4
    a_front: float = 0.0
    a_rear: float = 0.0
    b: float = 0.0
    c: float = 0.0
    depth_front: float = 2.0
    depth_rear: float = 4.0
    sigma_scale: float = 1.0

    def resolved_axes(self, twin) -> tuple[float, float, float, float]:
        sig = max(0.5, twin.sigma_cells * self.sigma_scale)
        a_f = self.a_front if self.a_front > 0.0 else sig
        a_r = self.a_rear if self.a_rear > 0.0 else max(sig, 2.0 * sig)
        b = self.b if self.b > 0.0 else sig
        c = self.c if self.c > 0.0 else max(
            0.5, self.depth_front if self.depth_front > 0.0 else self.depth_rear)
        return a_f, a_r, b, c

    def inject(self, twin, g, arc_i, arc_j, arc_k) -> None:
        sign = 1.0 if twin.travel_speed_m_s >= 0.0 else -1.0
        pen, en = _arc_weight_args(twin)
        a_f, a_r, b, c = self.resolved_axes(twin)
        kernels.inject_goldak_heat(g.H, g.flags, g.phi, g.f_l, g.arc_norm_buf, arc_i, arc_j,
                                   arc_k, twin.Q_w, g.dt, g.dx ** 3, twin.eta, sign, self.ff,
                                   self.fr, a_f, a_r, b, c, pen, en, g.FLAG_SOLID, g.FLAG_GAS,
                                   scratch=g.scratch)


@dataclass
class ConicalVolume:
    """Widened Gaussian centred 1.5 cells below the surface."""

    sigma_scale: float = 1.2

    def inject(self, twin, g, arc_i, arc_j, arc_k) -> None:
        pen, en = _arc_weight_args(twin)
        kernels.inject_arc_heat(g.H, g.flags, g.phi, g.f_l, g.arc_norm_buf, arc_i, arc_j,
                                arc_k - 1.5, twin.Q_w, twin.sigma_cells * self.sigma_scale, g.dt,
                                g.dx ** 3, twin.eta, pen, en, g.FLAG_SOLID, g.FLAG_GAS,
                                scratch=g.scratch)


def create_heat_source(name: str, **kwargs):
    key = (name or "gaussian2d").lower().replace("-", "").replace("_", "")
    if key in ("goldak", "goldak3d", "doubleellipsoid"):
        ff, fr = normalize_goldak_fractions(float(kwargs.pop("ff", 0.6)),
                                            float(kwargs.pop("fr", 1.4)))
        return Goldak3D(ff=ff, fr=fr, **kwargs)
    if key in ("conical", "conicalvolume", "cone"):
        return ConicalVolume(**kwargs)
    return Gaussian2D()


__all__ = ["Gaussian2D", "Goldak3D", "ConicalVolume", "create_heat_source",
           "normalize_goldak_fractions"]
