"""Coupled-step throughput on one MI355X: the 768x384x192 bead-on-plate step with 16-bit storage.

Workload (the largest single-GPU configuration of the metric): low-carbon-steel plate
768x384x192 cells (56.6 M) at dx = 0.2 mm (dt = 20 us), bottom 24 layers solid at 300 K with
seeded +/-0.5 K noise, argon above; Gaussian2D arc (eta 0.8, 4 kW, sigma 2 mm) moving +x at
8 mm/s with 6.5 m/min wire feed; flow-tier forces, convective and radiative losses, two-rate
central-moment (MRT) collision; fp32 arithmetic everywhere, populations stored as f - w_q in
binary16 (the value) and, in the same line, in fp32 and bfloat16 from the same seeded state.

One step = one full coupled timestep (reference solvers/coupled_step.py:73-464). The JSON line
reports MLUPS = nx*ny*nz*steps/s/1e6 over the timed steps (state resident in HBM), the fp32
run's MLUPS and the speed-up of the 16-bit forms, the roofline of the dominant kernel measured
with HIP events on its stream, the whole step against its compulsory bytes by cell class, and a
CPU baseline (the OpenMP build of the oracle and one thread of it, timed on a bounded sample of
the same workload on this host). ``--workload bead_on_plate`` runs the same physics on
256x128x64 cells.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# The bead-on-plate physics on 256 x 128 x 64 cells (51.2 x 25.6 x 12.8 mm at 0.2 mm).
WORKLOAD = dict(nx=256, ny=128, nz=64, dx=2e-4, plate_thickness_mm=4.8, arc_power_W=4000.0,
                arc_efficiency=0.8, arc_sigma_mm=2.0, travel_speed_m_s=8e-3, x0_mm=10.0,
                wire_feed_m_min=6.5, wire_diameter_mm=1.2, droplet_length_mm=3.0,
                welding_current_A=140.0, steps=4000)
# The same physics on 768 x 384 x 192 cells (56.6 M, 153.6 x 76.8 x 38.4 mm at 0.2 mm), the
# largest single-GPU configuration named for this path (fp32 state: about 16 GB of HBM). The
# timed steps start from the state after `preroll` steps of welding (untimed, captured steps):
# by then the plate under the arc has warmed and the first droplet (near step 2400) has been
# placed, so the window steps a deposited bead rather than the plate right after reset.
LARGE = dict(WORKLOAD, nx=768, ny=384, nz=192, x0_mm=30.0, steps=400, preroll=2600)
# Three passes on one line with interpass cooling, 20 000 steps on 768 x 384 x 192 cells at
# dx = 1 mm (dt = 0.1 ms: two 0.5 s interpasses take 10 000 steps at this dx; at 0.2 mm they
# alone would need 50 000): passes of 3334 welding steps (26.7 mm at 8 mm/s), interpasses of
# 4999 steps with the arc off at the end of the pass; 16-bit populations, captured steps.
MULTIPASS = dict(WORKLOAD, nx=768, ny=384, nz=192, dx=1e-3, plate_thickness_mm=48.0,
                 x0_mm=200.0, passes=3, pass_steps=3334, interpass_steps=4999)
# The reference README's job (jobs/examples/bead_on_plate.yaml: ER70S-6 with temperature-
# dependent cp, k, mu and Sahoo dgamma/dT, Goldak source, Lin-Eagar pressure, droplets, flow
# tier) on its own grid, 200 x 100 x 62 cells at 0.4 mm, stepped along its torch path.
JOB = dict(job="jobs/examples/bead_on_plate.yaml", steps=400)
WORKLOADS = {"bead_on_plate": WORKLOAD, "large": LARGE, "multipass": MULTIPASS, "job": JOB}


def twin_kwargs(w: dict) -> dict:
    return dict(nx=w["nx"], ny=w["ny"], nz=w["nz"], dx=w["dx"], arc_power_W=w["arc_power_W"],
                arc_efficiency=w["arc_efficiency"], arc_sigma_mm=w["arc_sigma_mm"],
                use_srt=False, enable_heat_loss=True, enable_convection=True,
                enable_radiation=True, enable_vof=True, enable_csf_tension=True,
                enable_wetting=True, enable_hydrostatic_gravity=True, enable_bead_freeze=True,
                enable_droplet_impact_pressure=True, max_tracers=20000,
                wire_diameter_mm=w["wire_diameter_mm"],
                droplet_freq_hz=(w["wire_feed_m_min"] / 60.0) / (w["droplet_length_mm"] / 1000.0))


def configure(twin, w: dict) -> None:
    twin.plate_thickness_mm = w["plate_thickness_mm"]
    twin.travel_speed_m_s = w["travel_speed_m_s"]
    twin.wire_feed_m_s = w["wire_feed_m_min"] / 60.0
    twin.welding_current_A = w["welding_current_A"]


def seed_noise(T, H, flags, cp_rho, seed=1234):
    import numpy as np

    rng = np.random.default_rng(seed)
    metal = flags != 2
    T[metal] = (T[metal] + rng.uniform(-0.5, 0.5, T.shape).astype(np.float32)[metal]).astype(np.float32)
    H[metal] = (np.float32(cp_rho) * T[metal]).astype(np.float32)


def torch_xy(w: dict, step: int, dt: float) -> tuple[float, float]:
    return (w["x0_mm"] * 1e-3 + w["travel_speed_m_s"] * step * dt,
            (w["ny"] // 2) * w["dx"])


def build_gpu_twin(w: dict, pop: str = "fp32"):
    from waam_solver_amd.twin import WAAMTwin

    tw = WAAMTwin(material="LCS-bead-on-plate", population_storage=pop, **twin_kwargs(w))
    configure(tw, w)
    tw.reset()
    g = tw.grid
    T, H, flags = g.T.to_numpy(), g.H.to_numpy(), g.flags.to_numpy()
    seed_noise(T, H, flags, tw.cp_rho)
    g.T.from_numpy(T)
    g.H.from_numpy(H)
    return tw


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads(n: int | None) -> None:
    """Set the OpenMP thread count of the loaded oracle (None: leave it)."""
    if n is None:
        return
    import ctypes

    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(w: dict, seconds: float, one_thread_steps: int) -> dict:
    """The OpenMP oracle on this host over a bounded number of steps of the same workload, then
    one thread of it over `one_thread_steps` further steps."""
    from oracle.oracle import Material, OracleTwin
    from waam_solver_amd.materials import load_material

    m = load_material("LCS-bead-on-plate")
    mat = Material(name=m.name, k=m.k, rho=m.rho, cp=m.cp, T_solidus=m.T_solidus,
                   T_liquidus=m.T_liquidus, L_fusion=m.L_fusion, gamma_0=m.gamma_0,
                   dgamma_dT=m.dgamma_dT, mu=m.mu, beta_T=m.beta_T, alpha=m.alpha)
    ot = OracleTwin(mat, parallel=True, **twin_kwargs(w))
    configure(ot, w)
    ot.reset()
    seed_noise(ot.g.T, ot.g.H, ot.g.flags, ot.cp_rho)
    ncell = w["nx"] * w["ny"] * w["nz"]
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    _omp_threads(threads)
    x, y = torch_xy(w, 0, ot.g.dt)
    ot.step(x, y, True)  # first step outside the timing (library load, page faults)
    steps, t0 = 0, time.perf_counter()
    while True:
        x, y = torch_xy(w, steps + 1, ot.g.dt)
        ot.step(x, y, True)
        steps += 1
        el = time.perf_counter() - t0
        if (el >= seconds and steps >= 2) or steps >= 200:
            break
    out = {"value": round(ncell * steps / el / 1e6, 3), "unit": "MLUPS", "cores": threads,
           "kind": "port", "cpu": cpu_model(),
           "sample": f"{steps} coupled steps of the same {w['nx']}x{w['ny']}x{w['nz']} workload "
                     f"(steps 2..{steps + 1}) on the OpenMP build of the oracle with {threads} "
                     f"threads, {el:.1f} s"}
    if one_thread_steps > 0:
        _omp_threads(1)
        t1 = time.perf_counter()
        for s in range(one_thread_steps):
            x, y = torch_xy(w, steps + 2 + s, ot.g.dt)
            ot.step(x, y, True)
        e1 = time.perf_counter() - t1
        _omp_threads(threads)
        out["one_thread"] = {"value": round(ncell * one_thread_steps / e1 / 1e6, 3),
                             "unit": "MLUPS", "cores": 1,
                             "sample": f"{one_thread_steps} further steps on one thread, "
                                       f"{e1:.1f} s"}
    return out


# The dominant kernel is k_step_pull, the deferred pull stream of the populations (a launch of
# its own ahead of k_step_stream). Its algorithmic bytes per launch (each array element touched
# once; neighbour re-reads not counted):
#   every cell        flags read, previous stream flags read                         2 B
#   pulled cell       19 populations read + 19 written: fp32 152 B, 16-bit storage  76 B
# A pulled cell is one the collision skips this step that was not SOLID at the step whose
# stream is being completed.
STREAM_B_EVERY = 2
STREAM_B_PULL = {"fp32": 152, "fp16": 76, "bf16": 76}
# Whole-step compulsory bytes by cell class (each persistent field read and written once):
#   every cell          H T f_l phi r/w 32 B, u r/w 24 B, flags r/w 2 B, rho w 4 B    62 B
#   non-SOLID cell      populations read + written: 2 x 19 x 4 B (fp32), 2 x 19 x 2 B (16-bit)
#   non-GAS cell        HAZ fields: T_max r/w, T_prev r/w, dT_dt w, three time-above r/w  44 B
# SOLID cells neither collide nor stream; GAS cells skip every HAZ update (reference
# kernels.py:2971-3010, 3016-3032, 2013-2027, 3261-3283).
STEP_B_EVERY = 62
STEP_B_POP = {"fp32": 152, "fp16": 76, "bf16": 76}
STEP_B_HAZ = 44
DOMINANT = "k_step_pull"
TIMING_EVERY = 4


def traffic_file(w: dict, pop: str) -> str:
    return os.path.join(ROOT, "profiles",
                        f"k_step_pull_bead_on_plate_{w['nx']}x{w['ny']}x{w['nz']}_{pop}_pmc.md")


def stream_classes(g) -> tuple[int, int]:
    """(collided cells, pulled cells) of the last fused step, from its stream flags."""
    cur = g.sflags.to_numpy()
    prev = g.sflags_tmp.to_numpy()
    active = (cur == 0) | (cur == 4)
    pull = (~active) & (prev != 1)
    return int(active.sum()), int(pull.sum())


def cell_classes(g) -> tuple[int, int]:
    """(SOLID cells, GAS cells) by the grid's current flags."""
    import torch

    from waam_solver_amd.field import tensor_of

    fl = tensor_of(g.flags)
    return int((fl == 1).sum().item()), int((fl == 2).sum().item())


def stream_bytes(ncell: int, pull: float, pop: str = "fp32") -> float:
    return STREAM_B_EVERY * ncell + STREAM_B_PULL[pop] * pull


def step_bytes(ncell: int, solid: float, gas: float, pop: str) -> float:
    return (STEP_B_EVERY * ncell + STEP_B_POP[pop] * (ncell - solid)
            + STEP_B_HAZ * (ncell - gas))


def timed_run(w: dict, args, pop: str):
    """Build the seeded twin with population storage `pop`, step the pre-roll (captured steps)
    and the warm-up, then time args.steps steps. Returns (twin, seconds, host seconds, (solid, gas) averaged over the
    start and end of the timed steps, (collided, pulled) likewise)."""
    import torch

    tw = build_gpu_twin(w, pop)
    tw.engine = args.engine
    g = tw.grid
    s0 = args.preroll
    if s0:
        pos = [torch_xy(w, s, g.dt) for s in range(s0)]
        tw.step_many([p[0] for p in pos], [p[1] for p in pos], True)
    for s in range(s0, s0 + args.warmup):
        tw.step(*torch_xy(w, s, g.dt), True)
    torch.cuda.synchronize()
    fused = args.engine == "fused"
    cls0 = cell_classes(g)
    sc0 = stream_classes(g) if fused and hasattr(g, "sflags_tmp") else None
    if fused:
        from waam_solver_amd._lib import STEP_KERNELS, lib

        # every fourth step's pull carries the timing events (each recorded dispatch costs a
        # few microseconds of the step)
        lib().call("waam_timing_begin_every", STEP_KERNELS.index(DOMINANT),
                   (args.steps + TIMING_EVERY - 1) // TIMING_EVERY, TIMING_EVERY)
    t0 = time.perf_counter()
    for s in range(s0 + args.warmup, s0 + args.warmup + args.steps):
        tw.step(*torch_xy(w, s, g.dt), True)
    t_host = time.perf_counter() - t0  # launches are asynchronous: the host's own cost per step
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    timing = None
    if fused:
        import ctypes

        tot_ms, n_launch = ctypes.c_double(), ctypes.c_int64()
        lib().call("waam_timing_end", ctypes.addressof(tot_ms), ctypes.addressof(n_launch))
        timing = (tot_ms.value, n_launch.value)
    cls1 = cell_classes(g)
    classes = ((cls0[0] + cls1[0]) / 2.0, (cls0[1] + cls1[1]) / 2.0)
    sc = None
    if fused:
        sc1 = stream_classes(g)
        sc0 = sc0 or sc1
        sc = ((sc0[0] + sc1[0]) / 2.0, (sc0[1] + sc1[1]) / 2.0)
    return tw, el, t_host, classes, sc, timing


def form_error(tw, ref) -> dict:
    """Relative L2 of T, phi, f_l and |u| of twin `tw` against twin `ref` (the same steps from
    the same state in another storage form), on the device."""
    import torch

    from waam_solver_amd.field import tensor_of

    out = {}
    for name in ("T", "phi", "f_l"):
        a = tensor_of(getattr(tw.grid, name)).double()
        b = tensor_of(getattr(ref.grid, name)).double()
        out[name] = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    ua = torch.sqrt(sum(tensor_of(getattr(tw.grid, c)).double() ** 2 for c in ("ux", "uy", "uz")))
    ub = torch.sqrt(sum(tensor_of(getattr(ref.grid, c)).double() ** 2 for c in ("ux", "uy", "uz")))
    out["u"] = float(torch.linalg.vector_norm(ua - ub) / torch.linalg.vector_norm(ub))
    return {k: float(f"{v:.3g}") for k, v in out.items()}


def compare_form(w: dict, args, pop: str):
    """The same workload, pre-roll, warm-up and timed steps with population storage `pop`:
    (twin, MLUPS, summary). The forms' errors against a binary64 build are measured over 2000
    steps with a molten pool by tests/precision_c3.py; main() adds each 16-bit form's error
    against this fp32 run over the same steps."""
    tw, el, _, classes, _, _ = timed_run(w, args, pop)
    g = tw.grid
    mlups = g.n_cells * args.steps / el / 1e6
    sb = step_bytes(g.n_cells, classes[0], classes[1], pop)
    out = {"value": round(mlups, 2), "ms_per_step": round(el * 1e3 / args.steps, 4),
           "step_bytes_per_cell": round(sb / g.n_cells, 1),
           "step_frac": round(sb * args.steps / el / 1e9 / HBM_PEAK_GBS, 4)}
    return tw, mlups, out


def kernel_breakdown(tw, w: dict, start: int, steps: int) -> dict:
    """Average HIP-event duration (us) of each fused-step kernel over `steps` further steps each
    (run after the timed region)."""
    import ctypes

    import torch

    from waam_solver_amd._lib import STEP_KERNELS, lib

    out, s = {}, start
    for kid, name in enumerate(STEP_KERNELS):
        lib().call("waam_timing_begin", kid, steps)
        for _ in range(steps):
            tw.step(*torch_xy(w, s, tw.grid.dt), True)
            s += 1
        torch.cuda.synchronize()
        tot, n = ctypes.c_double(), ctypes.c_int64()
        lib().call("waam_timing_end", ctypes.addressof(tot), ctypes.addressof(n))
        if n.value:
            out[name] = round(tot.value * 1e3 / n.value, 2)
    return out


def captured_form(tw, w: dict, start: int, steps: int) -> dict:
    """The same step count, but at least four graphs' worth, through WAAMTwin.step_many: runs of
    steady steps go to the GPU as captured graphs (droplet steps and the step after each go
    directly). Graphs are captured in an untimed first pass."""
    import torch

    g = tw.grid
    st = tw._fused
    warm = 4 * st.graph_steps
    steps = max(steps, 4 * st.graph_steps)
    pos = [torch_xy(w, s, g.dt) for s in range(start, start + warm + steps)]
    tw.step_many([p[0] for p in pos[:warm]], [p[1] for p in pos[:warm]], True)
    torch.cuda.synchronize()
    c0 = st.captured_steps
    t0 = time.perf_counter()
    tw.step_many([p[0] for p in pos[warm:]], [p[1] for p in pos[warm:]], True)
    t_host = time.perf_counter() - t0
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"value": round(g.n_cells * steps / el / 1e6, 2),
            "ms_per_step": round(el * 1e3 / steps, 4),
            "host_us_per_step": round(t_host * 1e6 / steps, 1),
            "steps_per_graph": st.graph_steps,
            "captured_steps": st.captured_steps - c0, "steps": steps}


def multipass(args) -> int:
    """The three-pass run: MLUPS over all 20 000 steps (droplet steps included), fp64 sums of
    rho over the metal cells and of the enthalpy (J) after every pass and interpass."""
    import torch

    from waam_solver_amd.field import tensor_of

    w = MULTIPASS
    pop = args.pop if args.pop != "fp32" or args.multipass_fp32 else "fp16"
    tw = build_gpu_twin(w, pop)
    g = tw.grid
    y0 = (g.ny // 2) * g.dx

    def sums(after: str) -> dict:
        tw.sync()
        metal = tensor_of(g.flags) != 2
        rho = float(tensor_of(g.rho).double()[metal].sum().item())
        h_j = float(tensor_of(g.H).double().sum().item()) * g.dx ** 3
        return {"after": after, "step": tw._step_n, "sum_rho_metal": rho, "sum_H_J": h_j,
                "metal_cells": int(metal.sum().item())}

    segs = []
    x_end = w["x0_mm"] * 1e-3
    for p in range(w["passes"]):
        xs = [w["x0_mm"] * 1e-3 + w["travel_speed_m_s"] * s * g.dt for s in range(w["pass_steps"])]
        segs.append((f"pass {p + 1}", xs, True))
        x_end = xs[-1]
        if p < w["passes"] - 1:
            segs.append((f"interpass {p + 1}", [x_end] * w["interpass_steps"], False))
    ledger = [sums("start")]
    el, steps = 0.0, 0
    for name, xs, weld in segs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tw.step_many(xs, [y0] * len(xs), weld)
        torch.cuda.synchronize()
        el += time.perf_counter() - t0
        steps += len(xs)
        ledger.append(sums(name))
        print(f"[multipass] {name}: {len(xs)} steps, {el:.1f} s so far", file=sys.stderr,
              flush=True)
    tel = tw.get_telemetry()
    t_weld = w["passes"] * w["pass_steps"] * g.dt
    first, last = ledger[0], ledger[-1]
    result = {
        "metric": f"coupled-step MLUPS (three-pass {g.nx}x{g.ny}x{g.nz} dx=1mm, {pop})",
        "value": round(g.n_cells * steps / el / 1e6, 2), "unit": "MLUPS", "n_gpus": 1,
        "steps": steps, "warmup": 0, "ms_per_step": round(el * 1e3 / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if pop == "fp32" else f"f32 arithmetic, {pop} populations",
        "data": "synthetic",
        "config": {"workload": f"three passes on one line, {w['pass_steps']} welding steps each "
                               f"(8 mm/s), {w['interpass_steps']}-step interpasses, "
                               f"{g.nx}x{g.ny}x{g.nz} dx=1mm dt=0.1ms, Gaussian2D 4kW eta0.8 "
                               "sigma2mm, wire feed 6.5m/min, flow tier, MRT",
                   "grid": [g.nx, g.ny, g.nz], "cells": g.n_cells,
                   "population_storage": pop, "captured_steps": tw._fused.captured_steps,
                   "steps_per_graph": tw._fused.graph_steps},
        "ledger": ledger,
        "drift": {"sum_rho_metal_rel": (last["sum_rho_metal"] - first["sum_rho_metal"])
                  / first["sum_rho_metal"],
                  "sum_H_rel": (last["sum_H_J"] - first["sum_H_J"]) / first["sum_H_J"],
                  "sum_H_gain_J": last["sum_H_J"] - first["sum_H_J"],
                  "arc_energy_J": w["arc_efficiency"] * w["arc_power_W"] * t_weld},
        "telemetry": {k: tel[k] for k in ("pool_width_mm", "pool_depth_mm", "peak_temp_K",
                                          "n_droplets_fired", "deposited_mass_g",
                                          "mass_balance_ratio") if k in tel},
    }
    print(json.dumps(result))
    return 0


def measured_traffic(w: dict, pop: str) -> dict | None:
    """HBM bytes per k_step_pull launch and per step from the committed rocprofv3 counter
    summary of this workload and population form."""
    out = {}
    try:
        with open(traffic_file(w, pop)) as fh:
            for line in fh:
                for key in ("traffic_bytes_per_launch", "traffic_bytes_per_step"):
                    if line.startswith(key + ":"):
                        out[key] = float(line.split(":", 1)[1])
    except OSError:
        return None
    return out or None


def job_run(args, engine: str) -> tuple[float, int, int]:
    """The README job through `engine` along its torch path (WAAMTwin.run_path for the warm-up
    steps, then the path's next positions through step_many, timed): (seconds, cells, droplets
    fired)."""
    import torch

    from waam_solver_amd.job import PKG_ROOT, load_job_config, parse_torch_path
    from waam_solver_amd.torch_path import TorchPathDriver, clamp_torch_to_domain
    from waam_solver_amd.twin import WAAMTwin

    path = PKG_ROOT / JOB["job"]
    tw = WAAMTwin.from_job(path)
    tw.engine = engine
    tw.reset()
    g = tw.grid
    job = load_job_config(path)
    tw.run_path(job, n_steps=args.warmup)
    driver = TorchPathDriver(parse_torch_path(job), tw.travel_speed_m_s)
    pos = list(driver.positions_for_steps(args.warmup + args.steps, g.dt))[args.warmup:]
    xy = [clamp_torch_to_domain(x, y, g.nx, g.ny, g.dx) for _, x, y, _ in pos]
    zs = [z for _, _, _, z in pos]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tw.step_many([c[0] for c in xy], [c[1] for c in xy], True, zs)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, g.n_cells, int(tw._n_droplets_fired)


def job_bench(args) -> int:
    """MLUPS of the README job with the fused step and with the reference-order kernels."""
    el, ncell, drops = job_run(args, "fused")
    el_ref, _, drops_ref = job_run(args, "reference")
    mlups = ncell * args.steps / el / 1e6
    ref = ncell * args.steps / el_ref / 1e6
    print(json.dumps({
        "metric": "coupled-step MLUPS (README job: bead_on_plate.yaml, table material)",
        "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "README job jobs/examples/bead_on_plate.yaml: ER70S-6 with "
                               "temperature-dependent tables (Sahoo dgamma/dT), Goldak, "
                               "Lin-Eagar pressure, droplets, flow tier, along its torch path",
                   "cells": ncell, "engine": "fused", "droplets": drops},
        "reference_order_engine": {"value": round(ref, 2), "ms_per_step":
                                   round(el_ref * 1e3 / args.steps, 4), "droplets": drops_ref},
        "speedup_vs_reference_order": round(mlups / ref, 3)}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=tuple(WORKLOADS), default="large",
                    help="large: 768x384x192 (default); bead_on_plate: 256x128x64; multipass: "
                         "three passes, 20 000 steps on 768x384x192 at dx=1mm; job: the README "
                         "job (table material) with the fused and the reference-order engine")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--preroll", type=int, default=None,
                    help="untimed captured steps before the warm-up (default: the workload's)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-one-thread-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", choices=("fused", "reference"), default="fused")
    ap.add_argument("--breakdown", type=int, default=20,
                    help="steps per kernel for the per-kernel HIP-event breakdown (0: skip)")
    ap.add_argument("--pop", choices=("fp32", "fp16", "bf16"), default="fp16",
                    help="population storage of the measured run (16-bit forms store f - w_q)")
    ap.add_argument("--multipass-fp32", action="store_true",
                    help="the multipass workload with fp32 populations (default: fp16)")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the captured-graph run reported beside the value")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the runs in the other population forms reported beside the value")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each box here has "
              "one GPU; multi-GPU sharding of the step is a design note, not a run)",
              file=sys.stderr)
        return 2

    import torch

    from waam_solver_amd._lib import STEP_KERNELS

    if args.workload == "multipass":
        return multipass(args)
    if args.workload == "job":
        if args.steps is None:
            args.steps = JOB["steps"]
        # one direct step and two 32-step captured runs: the graphs are captured before the
        # timed steps, which then replay them
        args.warmup = max(args.warmup, 65)
        return job_bench(args)
    w = dict(WORKLOADS[args.workload])
    if args.steps is None:
        args.steps = w["steps"]
    if args.preroll is None:
        args.preroll = w.get("preroll", 0)
    fused = args.engine == "fused"
    tw, el, t_host, classes, sc, timing = timed_run(w, args, args.pop)
    g = tw.grid
    ncell = g.nx * g.ny * g.nz
    mlups = ncell * args.steps / el / 1e6
    roofline = None
    if fused:
        tot_ms, n_launch = timing
        per_launch = stream_bytes(ncell, sc[1], args.pop)
        avg_s = tot_ms / 1e3 / max(n_launch, 1)
        achieved = per_launch / avg_s / 1e9
        meas = measured_traffic(w, args.pop)
        sb = step_bytes(ncell, classes[0], classes[1], args.pop)
        roofline = {"bound": "hbm", "kernel": DOMINANT, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": meas.get("traffic_bytes_per_launch") if meas else None,
                    "algorithmic_bytes_per_launch": round(per_launch),
                    "avg_launch_us": round(avg_s * 1e6, 2), "timed_launches": n_launch,
                    "timed_every": TIMING_EVERY,
                    "cells_collided": round(sc[0]), "cells_pulled": round(sc[1]),
                    "step": {"compulsory_bytes_per_cell": round(sb / ncell, 1),
                             "cells_solid": round(classes[0]), "cells_gas": round(classes[1]),
                             "achieved": round(sb * args.steps / el / 1e9, 1),
                             "frac": round(sb * args.steps / el / 1e9 / HBM_PEAK_GBS, 4),
                             "measured_bytes_per_step": (meas.get("traffic_bytes_per_step")
                                                         if meas else None)}}
    forms = None
    if fused and not args.no_precision:
        # the other population forms from the same seeded state over the same steps: fp32 (the
        # plain method) first, then the remaining 16-bit form
        forms = {}
        m32 = mlups
        ref32 = tw if args.pop == "fp32" else None
        for pop in ("fp32", "fp16", "bf16"):
            if pop == args.pop:
                continue
            other, m, out = compare_form(w, args, pop)
            if pop == "fp32":
                m32 = m
                ref32 = other
            else:
                out["speedup_vs_fp32"] = round(m / m32, 3)
                out["error_vs_fp32"] = form_error(other, ref32)
            forms[pop] = out
            if other is not ref32:
                del other
            torch.cuda.empty_cache()
        forms["value_speedup_vs_fp32"] = round(mlups / m32, 3)
        if ref32 is not tw:
            # the value's form against the fp32 run: the same pre-roll, warm-up and timed
            # steps from the same seeded state
            forms["value_error_vs_fp32"] = form_error(tw, ref32)
            forms["error_note"] = ("relative L2 against the fp32 run after the same steps; "
                                   "against a binary64 build see tests/precision_c3.py")
        del ref32
        torch.cuda.empty_cache()
    graph = None
    if fused:
        after = args.preroll + args.warmup + args.steps
        if args.breakdown > 0:
            roofline["kernel_us"] = kernel_breakdown(tw, w, after, args.breakdown)
            after += args.breakdown * len(STEP_KERNELS)
        if not args.no_graph:
            graph = captured_form(tw, w, after, args.steps)
    tel = tw.get_telemetry()
    result = {
        "metric": f"coupled-step MLUPS (bead-on-plate {g.nx}x{g.ny}x{g.nz}, {args.pop} "
                  "population storage)",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 4),
        "host_us_per_step": round(t_host * 1e6 / args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"bead-on-plate {g.nx}x{g.ny}x{g.nz} dx=0.2mm, Gaussian2D 4kW "
                               "eta0.8 sigma2mm +x 8mm/s, wire feed 6.5m/min, flow tier, MRT, "
                               f"{args.pop} populations (f - w_q), fp32 arithmetic; timed from "
                               f"the state after {args.preroll} untimed steps of welding",
                   "preroll_steps": args.preroll,
                   "grid": [g.nx, g.ny, g.nz], "cells": ncell, "collision": "mrt",
                   "engine": args.engine, "population_storage": args.pop,
                   "pool_width_mm": tel["pool_width_mm"], "pool_depth_mm": tel["pool_depth_mm"],
                   "peak_temp_K": tel["peak_temp_K"], "n_droplets": tel["n_droplets_fired"],
                   "surface_cells": round(sc[0]) if sc else None},
    }
    if roofline is not None:
        result["roofline"] = roofline
    if forms is not None:
        result["forms"] = forms
    if graph is not None:
        result["graph"] = graph
    if not args.no_cpu:
        del tw, g
        torch.cuda.empty_cache()
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds, args.cpu_one_thread_steps)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
