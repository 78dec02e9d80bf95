"""Thermocouple-style probes: T(t) and f_l(t) at world-fixed points (reference
export/probes.py:1-140).

A probe holds a WORLD position and is re-resolved to a cell at each sample, so it stays put when
the moving window shifts; outside the window it samples NaN. Sampling stays on the device: one
gather launch per step writes the probes' (T, f_l) into a row of a device ring, and the ring is
read back in one copy when the rows are needed (``rows``, ``write_csv``) or the ring fills — the
step loop never waits on a probe.
"""

from __future__ import annotations

import csv
import math
import pathlib
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import kernels
from ..field import LAYOUT_ZYX, tensor_of


@dataclass
class ProbeSpec:
    name: str
    x_m: float
    y_m: float
    z_m: float

    def resolve(self, twin) -> tuple[int, int, int] | None:
        """Cell of the probe in the current window, or None outside it (reference
        export/probes.py:29-36: truncation toward zero of the window-relative position)."""
        g = twin.grid
        i = int((self.x_m - twin._window_offset_x_m) / g.dx)
        j = int(self.y_m / g.dx)
        k = int(self.z_m / g.dx)
        if not (0 <= i < g.nx and 0 <= j < g.ny and 0 <= k < g.nz):
            return None
        return i, j, k


def _linear_cell(g, ijk) -> int:
    if ijk is None:
        return -1
    i, j, k = ijk
    if g.layout == LAYOUT_ZYX:
        return (k * g.ny + j) * g.nx + i
    return (i * g.ny + j) * g.nz + k


@dataclass
class ProbeRecorder:
    probes: list[ProbeSpec] = field(default_factory=list)
    ring_rows: int = 4096
    _rows: list[dict[str, float | str]] = field(default_factory=list)

    def __post_init__(self):
        self._ring = None        # device (ring_rows, 2 n_probes) float32
        self._cells = None       # device int64 (n_probes,)
        self._cells_key = None   # (window offset, n_probes) the cells were resolved for
        self._pending: list[tuple[int, float]] = []  # (step, sim_time_ms) of each ring row

    # ------------------------------------------------------------------ probes
    def add_grid(self, i: int, j: int, k: int, twin, name: str | None = None) -> None:
        """A probe at the centre of cell (i, j, k) of the current window."""
        g = twin.grid
        self.probes.append(ProbeSpec(name or f"probe_{len(self.probes)}",
                                     (i + 0.5) * g.dx + twin._window_offset_x_m,
                                     (j + 0.5) * g.dx, (k + 0.5) * g.dx))

    def add_world_mm(self, x_mm: float, y_mm: float, z_mm: float, twin,
                     name: str | None = None) -> None:
        self.probes.append(ProbeSpec(name or f"probe_{len(self.probes)}", x_mm / 1000.0,
                                     y_mm / 1000.0, z_mm / 1000.0))

    @classmethod
    def from_job_list(cls, entries: list[dict], twin) -> "ProbeRecorder":
        """The job's ``probes`` list: entries with i, j, k are cells, others world mm."""
        rec = cls()
        for e in entries:
            name = str(e.get("name", f"probe_{len(rec.probes)}"))
            if "i" in e and "j" in e and "k" in e:
                rec.add_grid(int(e["i"]), int(e["j"]), int(e["k"]), twin, name)
            else:
                rec.add_world_mm(float(e.get("x_mm", 0)), float(e.get("y_mm", 0)),
                                 float(e.get("z_mm", 0)), twin, name)
        return rec

    # ------------------------------------------------------------------ sampling
    def _device_cells(self, twin):
        g = twin.grid
        key = (twin._window_offset_x_m, len(self.probes), g.layout)
        if self._cells is None or self._cells_key != key:
            cells = [_linear_cell(g, p.resolve(twin)) for p in self.probes]
            self._cells = torch.tensor(cells, dtype=torch.int64, device=tensor_of(g.T).device)
            self._cells_key = key
        return self._cells

    def record_step(self, twin) -> None:
        if not self.probes:
            return
        g = twin.grid
        n = len(self.probes)
        if self._ring is None or self._ring.shape[1] != 2 * n:
            self.flush()
            self._ring = torch.empty((self.ring_rows, 2 * n), dtype=torch.float32,
                                     device=tensor_of(g.T).device)
        if len(self._pending) == self.ring_rows:
            self.flush()
        row = self._ring[len(self._pending)]
        kernels.probe_gather_into(g.T, g.f_l, self._device_cells(twin), row)
        self._pending.append((twin._step_n, round(twin._step_n * g.dt * 1000.0, 4)))

    def flush(self) -> None:
        """Move the device ring's rows to the host row list (one copy, one synchronisation)."""
        if not self._pending:
            return
        vals = self._ring[:len(self._pending)].cpu().numpy().astype(np.float64)
        for (step, t_ms), v in zip(self._pending, vals):
            row: dict[str, float | str] = {"sim_time_ms": t_ms, "step": step}
            for p, (T, fl) in zip(self.probes, v.reshape(-1, 2)):
                if math.isnan(T):
                    row[f"{p.name}_T_K"] = math.nan
                    row[f"{p.name}_f_l"] = math.nan
                else:
                    row[f"{p.name}_T_K"] = round(float(np.float32(T)), 2)
                    row[f"{p.name}_f_l"] = round(float(np.float32(fl)), 4)
            self._rows.append(row)
        self._pending.clear()

    @property
    def rows(self) -> list[dict[str, float | str]]:
        self.flush()
        return self._rows

    def write_csv(self, path: str | pathlib.Path) -> None:
        """sim_time_ms, step, then <name>_T_K and <name>_f_l per probe (reference
        export/probes.py:115-129)."""
        rows = self.rows
        if not rows:
            return
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        names = ["sim_time_ms", "step"]
        for p in self.probes:
            names += [f"{p.name}_T_K", f"{p.name}_f_l"]
        with open(path, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=names, extrasaction="ignore")
            w.writeheader()
            w.writerows(rows)

    def clear(self) -> None:
        self._pending.clear()
        self._rows.clear()
