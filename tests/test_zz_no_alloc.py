"""The allocation-free step (waam_step_fused_noalloc) against the ordinary one (waam_step_fused).

Two twins per grid start from the same seeded state; one steps through waam_step_fused with the
grid's own workspace and cell lists, the other through waam_step_fused_noalloc, which uses the
library's device work array instead. Every field and both population buffers must be equal bit
for bit. The calls interleave: a step on the 64x40x32 grid is followed at once by a step on a
smaller 48x32x24 grid (the work array keeps the larger grid's contents), then each grid takes
further steps, the first of them with its thermal list rebuilt as the entry's contract asks.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.step_cases import FIELDS, build_pair, torch_x

pytestmark = pytest.mark.gpu

NOALLOC = "waam_step_fused_noalloc"


def _twins(**overrides):
    from waam_solver_amd.solvers.fused import FusedStepper

    a, _ = build_pair("bead_mrt", **overrides)
    b, _ = build_pair("bead_mrt", **overrides)
    b._fused = FusedStepper(b)
    b._fused.entry = NOALLOC
    return a, b


def _step(tw, s: int) -> None:
    g = tw.grid
    tw.step(torch_x(g.dx, g.nx, s, g.dt), (g.ny // 2) * g.dx, True)


def _differing(a, b) -> list[str]:
    a.sync()
    b.sync()
    ga, gb = a.grid, b.grid
    bad = []
    for name in FIELDS + ("f_src", "f_dst"):
        x, y = getattr(ga, name).to_numpy(), getattr(gb, name).to_numpy()
        if x.tobytes() != y.tobytes():
            bad.append(name)
    return bad


def test_noalloc_step_matches_the_workspace_step():
    import torch

    print("[noalloc] bead_mrt 64x40x32 then 48x32x24", flush=True)
    big_a, big_b = _twins()
    small_a, small_b = _twins(nx=48, ny=32, nz=24)
    H0 = big_b.grid.H.to_numpy()
    # two calls in a row on one GPU, the second on the smaller grid
    _step(big_a, 0)
    _step(big_b, 0)
    _step(small_a, 0)
    _step(small_b, 0)
    # the work array now holds the small grid's lists: the big grid's next call rebuilds its own
    big_b._fused._list_epoch = None
    for s in range(1, 4):
        _step(big_a, s)
        _step(big_b, s)
    small_b._fused._list_epoch = None
    for s in range(1, 4):
        _step(small_a, s)
        _step(small_b, s)
    torch.cuda.synchronize()
    assert big_b._fused.entry == NOALLOC and small_b._fused.entry == NOALLOC
    assert not _differing(big_a, big_b), "64x40x32"
    assert not _differing(small_a, small_b), "48x32x24"
    # the runs did something: the arc heated the plate
    assert float(np.max(big_b.grid.H.to_numpy() - H0)) > 0.0


def test_noalloc_rejects_grids_beyond_its_array():
    from waam_solver_amd._lib import WaamError, lib
    from waam_solver_amd.solvers.fused import StepFields, StepParams

    L = lib()
    cap = int(L.dll.waam_noalloc_max_cells())
    f = StepFields()
    f.nx, f.ny, f.nz = 1024, 1024, cap // (1024 * 1024) + 1
    p = StepParams()
    with pytest.raises(WaamError, match="device work array"):
        L.call(NOALLOC, ctypes.addressof(p), ctypes.addressof(f), None)
