"""Coupled-step cases shared by the GPU parity tests and the smoke check.

A case builds a product twin (GPU) and an oracle twin (CPU) from the same settings, applies the
same host-generated fixture to both (seeded 1 K temperature noise on the metal and a seeded
molten patch, reference tools/force_ablation.py:71-111) and drives both along the same torch
path. The bead-on-plate settings follow the 256x128x64 workload: low-carbon-steel substrate,
Gaussian2D arc (eta 0.8, 4 kW), flow-tier forces, convective and radiative losses.
"""

from __future__ import annotations

import numpy as np

from tests.fixtures import ER70S6_V1_TABLES, LCS_BEAD_ON_PLATE, SOLID, seed_pool

FLOW_TIER = dict(enable_vof=True, enable_csf_tension=True, enable_wetting=True,
                 enable_hydrostatic_gravity=True, enable_bead_freeze=True,
                 enable_droplet_impact_pressure=True)


def case_kwargs(name: str) -> dict:
    base = dict(arc_power_W=4000.0, arc_efficiency=0.8, arc_sigma_mm=2.0, use_srt=False,
                enable_heat_loss=True, enable_convection=True, enable_radiation=True,
                max_tracers=2000, **FLOW_TIER)
    if name == "bead_mrt":
        return dict(base, nx=64, ny=40, nz=32, dx=2e-4, droplet_freq_hz=0.0)
    if name == "droplets_srt":
        return dict(base, nx=56, ny=40, nz=32, dx=2e-4, use_srt=True, droplet_freq_hz=2500.0)
    if name == "tables_mrt":
        # temperature-dependent cp, k, mu, dgamma/dT: variable-tau SRT collision, per-cell cp
        return dict(base, nx=64, ny=40, nz=32, dx=2e-4, droplet_freq_hz=0.0)
    if name == "bench_small":
        return dict(base, nx=128, ny=64, nz=32, dx=2e-4, droplet_freq_hz=0.0)
    raise KeyError(name)


def configure(twin, name: str) -> None:
    twin.plate_thickness_mm = 1.6
    twin.travel_speed_m_s = 8e-3
    if name == "droplets_srt":
        twin.wire_feed_m_s = 6.5 / 60.0
        twin.welding_current_A = 140.0


def fixture_arrays(T, f_l, phi, flags, H, nz_solid, mat, cp_rho, L_rho, seed: int = 7):
    """Seeded +/-0.5 K noise on metal (H = cp_rho*T) and a molten patch of radius 5."""
    rng = np.random.default_rng(seed)
    metal = flags != 2
    noise = rng.uniform(-0.5, 0.5, T.shape).astype(np.float32)
    T[metal] = (T[metal] + noise[metal]).astype(np.float32)
    H[metal] = (np.float32(cp_rho) * T[metal]).astype(np.float32)
    seed_pool(T, f_l, phi, flags, H, nz_solid, mat["T_liquidus"], cp_rho, L_rho, radius=5)


def torch_x(twin_dx: float, nx: int, step: int, dt: float) -> float:
    return (nx // 3) * twin_dx + 8e-3 * step * dt


def build_pair(name: str, mode: str = "jacobi", engine: str = "fused", layout: int | None = None):
    """(product twin, oracle twin), both reset and seeded identically."""
    from oracle.oracle import Material, OracleTwin
    from waam_solver_amd.materials import material_from_dict
    from waam_solver_amd.twin import WAAMTwin

    kw = case_kwargs(name)
    mat_d = dict(LCS_BEAD_ON_PLATE)
    tables = ER70S6_V1_TABLES if name.startswith("tables") else {}
    prod_mat = material_from_dict({"name": mat_d["name"], "status": "synthetic",
                                   "constants": {k: v for k, v in mat_d.items() if k != "name"},
                                   "tables": {k: [list(p) for p in v] for k, v in tables.items()}})
    tw = WAAMTwin(material=prod_mat, field_layout=layout, **kw)
    tw.engine = engine
    ot = OracleTwin(Material(**mat_d, tables=tables), mode=mode, **kw)
    assert tw.use_material_tables == ot.use_tables == bool(tables)
    for t in (tw, ot):
        configure(t, name)
    tw.reset()
    ot.reset()
    g, og = tw.grid, ot.g
    T, f_l, phi = og.T.copy(), og.f_l.copy(), og.phi.copy()
    flags, H = og.flags.copy(), og.H.copy()
    fixture_arrays(T, f_l, phi, flags, H, ot.nz_solid, mat_d, ot.cp_rho, ot.L_rho)
    for arr_name, arr in (("T", T), ("f_l", f_l), ("phi", phi), ("flags", flags), ("H", H)):
        getattr(og, arr_name)[...] = arr
        getattr(g, arr_name).from_numpy(arr)
    return tw, ot


def run_pair(tw, ot, n_steps: int, start: int = 0) -> None:
    g = tw.grid
    for s in range(start, start + n_steps):
        x = torch_x(g.dx, g.nx, s, g.dt)
        y = (g.ny // 2) * g.dx
        tw.step(x, y, True)
        ot.step(x, y, True)


FIELDS = ("H", "T", "f_l", "phi", "flags", "rho", "ux", "uy", "uz", "T_max", "T_prev", "dT_dt",
          "Fx", "Fy", "Fz", "time_above_solidus_s")


def compare(tw, ot, fields=FIELDS) -> dict:
    """{field: (bit_equal, rel_l2)} plus the two population buffers."""
    from tests.gpu_common import bits_equal, rel_l2

    out = {}
    tw.sync()
    g, og = tw.grid, ot.g
    for f in fields:
        a = getattr(g, f).to_numpy()
        b = getattr(og, f)
        if f == "flags":
            a = a.astype(np.uint8)
        out[f] = (bits_equal(a, b), 0.0 if f == "flags" else rel_l2(a, b))
    for name, a, b in (("f_src", g.f_src.to_numpy(), og.f_src), ("f_dst", g.f_dst.to_numpy(), og.f_dst)):
        out[name] = (bits_equal(a, b), rel_l2(a, b))
    return out
