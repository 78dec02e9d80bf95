"""Captured steps (waam_graph_capture / waam_graph_launch) against direct steps and the oracle.

A run of steps through ``WAAMTwin.step_many`` sends runs of steps to the GPU as one captured
graph each, droplet steps included (their deposition runs on the device with the parameters
recorded into the graph); every field and both population buffers must equal, bit for bit,
those of the oracle twin stepped one step at a time on the same torch path (the direct fused
step is held to the same bar by the step parity tests).
"""

from __future__ import annotations

import pytest

from tests.step_cases import FIELDS, build_pair, compare, torch_x

pytestmark = pytest.mark.gpu


def _positions(g, start: int, n: int):
    xs = [torch_x(g.dx, g.nx, s, g.dt) for s in range(start, start + n)]
    return xs, [(g.ny // 2) * g.dx] * n


@pytest.mark.parametrize("name,pop,steps,per_graph", [
    ("bead_mrt", "fp32", 40, 8),
    ("surface_forces_mrt", "fp32", 33, 16),
    ("droplets_srt", "fp32", 61, 6),
    ("droplets_srt", "fp16", 61, 6),
    ("bead_mrt", "fp16", 40, 8),
    ("bead_odd_mrt", "bf16", 24, 7),
    ("tables_mrt", "fp32", 40, 8),
    ("tables_mrt", "fp16", 24, 8),
])
def test_captured_steps_match_oracle(name, pop, steps, per_graph):
    print(f"[graph] {name} {pop} steps={steps} per_graph={per_graph}", flush=True)
    tw, ot = build_pair(name, population_storage=pop)
    g = tw.grid
    # one direct step first: a captured run follows a fused step
    xs, ys = _positions(g, 0, steps)
    tw.step(xs[0], ys[0], True)
    tw._fused.graph_steps = per_graph
    tw.step_many(xs[1:], ys[1:], True)
    for x, y in zip(xs, ys):
        ot.step(x, y, True)
    assert tw._step_n == ot._step_n == steps
    captured = tw._fused.captured_steps
    assert captured >= per_graph, "no captured run"
    if name == "tables_mrt":
        # the property fields of the last refresh too
        res = compare(tw, ot, FIELDS + ("cp_rho_field", "alpha_lu_field", "dgamma_lu_field",
                                        "tau_field"))
    else:
        res = compare(tw, ot)
    bad = {k: v for k, v in res.items() if not v[0]}
    assert not bad, f"{name} {pop}: captured steps differ from the oracle: {bad}"
    if name == "droplets_srt":
        # the droplet steps ran inside the graphs too (waam_step_fused_droplet captured with
        # its deposition parameters): every step after the first was captured
        assert captured == steps - 1, captured
        assert tw._n_droplets_fired == ot._n_droplets_fired > 0
        assert tw._deposited_volume_m3 == ot._deposited_volume_m3 > 0.0
        assert tw._deposition_overflow == ot._deposition_overflow
    assert len(tw._fused._graphs) >= 1
    tw._fused.close()


def test_graph_relaunch_after_direct_steps():
    """Graphs are kept and launched again after direct steps, a sync and a host write in
    between (the write forces one direct step before the next captured run)."""
    print("[graph] relaunch bead_mrt", flush=True)
    tw, ot = build_pair("bead_mrt")
    g = tw.grid
    xs, ys = _positions(g, 0, 60)
    tw.step(xs[0], ys[0], True)
    tw._fused.graph_steps = 10
    tw.step_many(xs[1:21], ys[1:21], True)
    tw.sync()
    t = g.T.to_numpy()
    g.T.from_numpy(t)  # a host write: the next step must build its own thermal list
    tw.step_many(xs[21:60], ys[21:60], True)
    for x, y in zip(xs, ys):
        ot.step(x, y, True)
    assert tw._fused.captured_steps >= 30
    res = compare(tw, ot)
    bad = {k: v for k, v in res.items() if not v[0]}
    assert not bad, f"relaunch: {bad}"
