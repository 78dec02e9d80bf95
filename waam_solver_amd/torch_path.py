"""Torch path interpolation (reference torch_path.py)."""

from __future__ import annotations

import csv
import math
import pathlib
from dataclasses import dataclass
from typing import Iterator


@dataclass
class PathSegment:
    x0: float
    y0: float
    z0: float
    x1: float
    y1: float
    z1: float
    length_m: float
    cumulative_start_m: float = 0.0


def build_segments(waypoints) -> list[PathSegment]:
    segs, cum = [], 0.0
    for (x0, y0, z0), (x1, y1, z1) in zip(waypoints[:-1], waypoints[1:]):
        length = math.sqrt((x1 - x0) ** 2 + (y1 - y0) ** 2 + (z1 - z0) ** 2)
        if length < 1e-9:
            continue
        segs.append(PathSegment(x0, y0, z0, x1, y1, z1, length, cum))
        cum += length
    return segs


def load_torch_path_csv(path) -> list[tuple[float, float, float]]:
    out = []
    with open(pathlib.Path(path), newline="") as f:
        for row in csv.DictReader(f):
            out.append(tuple(float(row.get(f"{a}_mm", row.get(a, 0))) / 1000.0 for a in "xyz"))
    return out


def clamp_torch_to_domain(x_m, y_m, nx, ny, dx, margin_cells=4):
    return (max(margin_cells * dx, min(x_m, (nx - margin_cells) * dx)),
            max(margin_cells * dx, min(y_m, (ny - margin_cells) * dx)))


class TorchPathDriver:
    """Piecewise-linear waypoints traversed at constant speed; parks at the end of the path."""

    def __init__(self, waypoints, travel_speed_m_s: float):
        self.speed = max(travel_speed_m_s, 1e-9)
        self.segments = build_segments(list(waypoints))
        self.total_length = sum(s.length_m for s in self.segments)
        self._fallback = waypoints[0] if (not self.segments and waypoints) else (0.0, 0.0, 0.0)

    def distance_at_step(self, step: int, dt: float) -> float:
        return step * dt * self.speed

    def segment_index_at_distance(self, dist_m: float) -> int:
        if not self.segments:
            return 0
        d = min(dist_m, self.total_length)
        for idx, seg in enumerate(self.segments):
            if d <= seg.length_m:
                return idx
            d -= seg.length_m
        return len(self.segments) - 1

    def segment_end(self, idx: int):
        if not self.segments or idx < 0 or idx >= len(self.segments):
            return None
        s = self.segments[idx]
        return s.x1, s.y1, s.z1

    def position_at_time(self, t: float):
        if not self.segments:
            return self._fallback
        dist = min(t * self.speed, self.total_length)
        for s in self.segments:
            if dist <= s.length_m:
                f = dist / s.length_m
                return s.x0 + f * (s.x1 - s.x0), s.y0 + f * (s.y1 - s.y0), s.z0 + f * (s.z1 - s.z0)
            dist -= s.length_m
        last = self.segments[-1]
        return last.x1, last.y1, last.z1

    def positions_for_steps(self, n_steps: int, dt: float, start_step: int = 0) -> Iterator:
        for step in range(start_step, start_step + n_steps):
            x, y, z = self.position_at_time(step * dt)
            yield step, x, y, z
