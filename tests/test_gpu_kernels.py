"""Per-kernel parity: every HIP entry point against the oracle, bit for bit, on seeded states.

Each case builds a seeded random state (plate, gas, liquid patch, interface band and sprinkled
random flags so every branch runs), hands identical bytes to the GPU kernel and to the oracle's
restatement of the same reference kernel, and requires identical output bits (fp32 fields,
uint8 flags). Both sides evaluate fp32 expressions in the reference's operand order with FMA
contraction off and the shared exp(), so any difference is a defect.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from tests.gpu_common import bits_equal, mismatch_report, random_state

pytestmark = pytest.mark.gpu

DIMS = [(24, 20, 16), (17, 13, 70)]


@pytest.fixture(scope="module")
def env(hip_device):
    import torch

    from oracle.oracle import lib as olib
    from waam_solver_amd import kernels as K

    return dict(torch=torch, K=K, O=olib(), dev=hip_device)


def up(env, a):
    return env["torch"].from_numpy(np.ascontiguousarray(a)).to(env["dev"])


def down(t):
    return t.detach().cpu().numpy()


def P(a):
    from oracle.oracle import ptr

    return ptr(a)


def check(name, gpu, ora):
    assert bits_equal(gpu, ora), mismatch_report(name, gpu, ora)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("which", ["mrt", "srt"])
def test_collide(env, dims, which):
    s = random_state(*dims, seed=11)
    n = dims
    g = {k: up(env, s[k]) for k in ("f", "f2", "rho", "ux", "uy", "uz", "Fx", "Fy", "Fz", "f_l", "flags")}
    o = {k: s[k].copy() for k in g}
    K, O = env["K"], env["O"]
    if which == "mrt":
        K.collide_mrt(g["f"], g["f2"], g["rho"], g["ux"], g["uy"], g["uz"], g["Fx"], g["Fy"], g["Fz"],
                      g["f_l"], g["flags"], None, None, None, None, None, 1.7, 1.3, 1.6e5, 1, 2, *n)
        O.or_collide_mrt(P(o["f"]), P(o["f2"]), P(o["rho"]), P(o["ux"]), P(o["uy"]), P(o["uz"]),
                         P(o["Fx"]), P(o["Fy"]), P(o["Fz"]), P(o["f_l"]), P(o["flags"]), 1.7, 1.3,
                         1.6e5, *n)
    else:
        K.collide_srt(g["f"], g["f2"], g["rho"], g["ux"], g["uy"], g["uz"], g["Fx"], g["Fy"], g["Fz"],
                      g["f_l"], g["flags"], 0.6, 1.0 / 0.6, 1.0, 1.6e5, 1, 2, *n)
        O.or_collide_srt(P(o["f"]), P(o["f2"]), P(o["rho"]), P(o["ux"]), P(o["uy"]), P(o["uz"]),
                         P(o["Fx"]), P(o["Fy"]), P(o["Fz"]), P(o["f_l"]), P(o["flags"]), 0.6,
                         1.0 / 0.6, 1.0, 1.6e5, *n)
    for k in ("f2", "rho", "ux", "uy", "uz"):
        check(k, down(g[k]), o[k])


@pytest.mark.parametrize("dims", DIMS)
def test_stream(env, dims):
    s = random_state(*dims, seed=12)
    g = {k: up(env, s[k]) for k in ("f", "f2", "flags")}
    o = {k: s[k].copy() for k in g}
    env["K"].stream(g["f"], g["f2"], g["flags"], 1, 2, *dims)
    env["O"].or_stream(P(o["f"]), P(o["f2"]), P(o["flags"]), *dims)
    check("f_dst", down(g["f2"]), o["f2"])


@pytest.mark.parametrize("dims", DIMS)
def test_thermal_advect_diffuse(env, dims):
    s = random_state(*dims, seed=13)
    g = {k: up(env, s[k]) for k in ("H", "T", "ux", "uy", "uz", "flags")}
    o = {k: s[k].copy() for k in g}
    alpha, cp = 0.05, 5.3e6
    env["K"].advect_diffuse_temperature(g["H"], g["T"], g["ux"], g["uy"], g["uz"], g["flags"],
                                        alpha, 1.0, 1, 2, cp, *dims)
    H_in = o["H"].copy()
    env["O"].or_advect_diffuse_temperature_jacobi(P(H_in), P(o["H"]), P(o["T"]), P(o["ux"]),
                                                  P(o["uy"]), P(o["uz"]), P(o["flags"]), alpha,
                                                  1.0, cp, *dims)
    check("H", down(g["H"]), o["H"])


@pytest.mark.parametrize("dims", DIMS)
def test_boundary_losses_floor_ceiling_phase(env, dims):
    s = random_state(*dims, seed=14)
    g = {k: up(env, s[k]) for k in ("H", "T", "flags", "f_l")}
    o = {k: s[k].copy() for k in g}
    K, O = env["K"], env["O"]
    cp, L, Ts, Tl = 5.3e6, 2.1e9, 1723.0, 1780.0
    K.apply_thermal_boundary_losses(g["H"], g["T"], g["flags"], 300.0, 25.0, 0.3, 1, 1, cp, 2e-5,
                                    2e-4, 5.670374419e-8, 1, 2, *dims)
    O.or_apply_thermal_boundary_losses(P(o["H"]), P(o["T"]), P(o["flags"]), 300.0, 25.0, 0.3, 1,
                                       1, cp, 2e-5, 2e-4, 5.670374419e-8, *dims)
    check("H losses", down(g["H"]), o["H"])
    K.clamp_enthalpy_floor_scalar(g["H"], cp, g["flags"], 1200.0, 2)
    O.or_clamp_enthalpy_floor_scalar(P(o["H"]), cp, P(o["flags"]), 1200.0, *dims)
    check("H floor", down(g["H"]), o["H"])
    K.clamp_enthalpy_ceiling_scalar(g["H"], g["flags"], cp, Ts, Tl, 2500.0, L, 2)
    O.or_clamp_enthalpy_ceiling_scalar(P(o["H"]), P(o["flags"]), cp, Ts, Tl, 2500.0, L, *dims)
    check("H ceiling", down(g["H"]), o["H"])
    K.update_phase(g["H"], g["T"], g["f_l"], cp, L, Ts, Tl)
    O.or_update_phase(P(o["H"]), P(o["T"]), P(o["f_l"]), cp, L, Ts, Tl, *dims)
    check("T", down(g["T"]), o["T"])
    check("f_l", down(g["f_l"]), o["f_l"])


def test_haz_diagnostics(env):
    dims = DIMS[0]
    s = random_state(*dims, seed=15)
    rng = np.random.default_rng(1)
    extra = {k: rng.uniform(0, 3000, dims).astype(np.float32) for k in ("Tm", "Tp", "d", "a", "b", "c")}
    g = {k: up(env, v) for k, v in {**s, **extra}.items() if isinstance(v, np.ndarray) and k != "f" and k != "f2"}
    o = {k: v.copy() for k, v in {**s, **extra}.items() if isinstance(v, np.ndarray)}
    K, O = env["K"], env["O"]
    K.update_T_max(g["T"], g["Tm"], g["flags"], 2)
    O.or_update_T_max(P(o["T"]), P(o["Tm"]), P(o["flags"]), *dims)
    K.update_cooling_rate(g["T"], g["Tp"], g["d"], g["flags"], 2e-5, 2)
    O.or_update_cooling_rate(P(o["T"]), P(o["Tp"]), P(o["d"]), P(o["flags"]), 2e-5, *dims)
    K.update_time_above_T(g["T"], g["flags"], g["a"], g["b"], g["c"], 2e-5, 1073.15, 1373.15,
                          1723.0, 2)
    O.or_update_time_above_T(P(o["T"]), P(o["flags"]), P(o["a"]), P(o["b"]), P(o["c"]), 2e-5,
                             1073.15, 1373.15, 1723.0, *dims)
    for k in ("Tm", "Tp", "d", "a", "b", "c"):
        check(k, down(g[k]), o[k])


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("source", ["gauss", "goldak"])
def test_arc_heat(env, dims, source):
    s = random_state(*dims, seed=16)
    g = {k: up(env, s[k]) for k in ("H", "flags", "f_l")}
    o = {k: s[k].copy() for k in g}
    ai, aj, ak = dims[0] * 0.45, dims[1] * 0.55, float(s["nz_solid"])
    norm_g = env["torch"].zeros((), dtype=env["torch"].float32, device=env["dev"])
    norm_o = np.zeros(1, np.float32)
    if source == "gauss":
        env["K"].inject_arc_heat(g["H"], g["flags"], None, g["f_l"], norm_g, ai, aj, ak, 4000.0,
                                 6.0, 2e-5, 8e-12, 0.8, 10.0, 1, 1, 2)
        env["O"].or_inject_arc_heat(P(o["H"]), P(o["flags"]), P(o["f_l"]), P(norm_o), ai, aj, ak,
                                    4000.0, 6.0, 2e-5, 8e-12, 0.8, 10.0, 1, *dims)
    else:
        env["K"].inject_goldak_heat(g["H"], g["flags"], None, g["f_l"], norm_g, ai, aj, ak,
                                    4000.0, 2e-5, 8e-12, 0.8, 1.0, 0.6, 1.4, 4.0, 8.0, 4.0, 3.0,
                                    10.0, 1, 1, 2)
        env["O"].or_inject_goldak_heat(P(o["H"]), P(o["flags"]), P(o["f_l"]), P(norm_o), ai, aj,
                                       ak, 4000.0, 2e-5, 8e-12, 0.8, 1.0, 0.6, 1.4, 4.0, 8.0, 4.0,
                                       3.0, 10.0, 1, *dims)
    assert float(norm_g.item()) == float(norm_o[0])
    check("H", down(g["H"]), o["H"])


@pytest.mark.parametrize("dims", DIMS)
def test_vof_chain(env, dims):
    s = random_state(*dims, seed=17)
    g = {k: up(env, s[k]) for k in ("phi", "ux", "uy", "uz", "flags", "f_l")}
    o = {k: s[k].copy() for k in g}
    K, O = env["K"], env["O"]
    torch = env["torch"]
    tmp_g = torch.empty_like(g["phi"])
    tmp_o = np.empty_like(o["phi"])
    K.advect_phi(tmp_g, g["phi"], g["ux"], g["uy"], g["uz"], g["flags"], 1, 2, *dims)
    O.or_advect_phi(P(tmp_o), P(o["phi"]), P(o["ux"]), P(o["uy"]), P(o["uz"]), P(o["flags"]), *dims)
    check("phi advect", down(tmp_g), tmp_o)
    K.reinitialize_phi(tmp_g, g["flags"], 1, 2, 0)
    O.or_reinitialize_phi(P(tmp_o), P(o["flags"]), *dims)
    check("phi reinit", down(tmp_g), tmp_o)
    theta = math.radians(80.0)
    K.apply_contact_angle_phi_bc_into(tmp_g, g["phi"], g["flags"], theta, *dims)
    c, sn = K.contact_angle_cos_sin(theta)
    O.or_apply_contact_angle_phi_bc_jacobi(P(tmp_o), P(o["phi"]), P(o["flags"]), c, sn, *dims)
    check("phi contact", down(g["phi"]), o["phi"])
    K.update_flags_from_phi(g["phi"], g["f_l"], g["flags"], s["nz_solid"], 0, 1, 2, 4)
    O.or_update_flags_from_phi(P(o["phi"]), P(o["f_l"]), P(o["flags"]), s["nz_solid"], *dims)
    check("flags", down(g["flags"]), o["flags"])
    K.sync_phi_liquid_fraction(g["phi"], g["f_l"], g["flags"], 2, 1)
    O.or_sync_phi_liquid_fraction(P(o["phi"]), P(o["f_l"]), P(o["flags"]), *dims)
    check("f_l", down(g["f_l"]), o["f_l"])


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("wetting", [0, 1])
def test_forces(env, dims, wetting):
    s = random_state(*dims, seed=18)
    g = {k: up(env, s[k]) for k in ("phi", "flags", "Fx", "Fy", "Fz", "T", "f_l")}
    o = {k: s[k].copy() for k in g}
    K, O = env["K"], env["O"]
    theta = math.radians(70.0)
    c, sn = K.contact_angle_cos_sin(theta)
    K.compute_csf_tension(g["phi"], g["flags"], g["Fx"], g["Fy"], g["Fz"], 3e-3, 1, 2, *dims,
                          wetting, theta)
    O.or_compute_csf_tension(P(o["phi"]), P(o["flags"]), P(o["Fx"]), P(o["Fy"]), P(o["Fz"]), 3e-3,
                             wetting, c, sn, *dims)
    K.compute_marangoni_force(g["T"], g["phi"], g["f_l"], g["Fx"], g["Fy"], g["Fz"], g["flags"],
                              -2e-7, 2e-4, 1, 2, *dims)
    O.or_compute_marangoni_force(P(o["T"]), P(o["phi"]), P(o["f_l"]), P(o["Fx"]), P(o["Fy"]),
                                 P(o["Fz"]), P(o["flags"]), -2e-7, *dims)
    ai, aj, ak = dims[0] * 0.5, dims[1] * 0.5, float(s["nz_solid"])
    K.apply_arc_pressure(g["Fz"], g["flags"], g["phi"], ai, aj, ak, 6.0, 500.0, 2e-5, 2e-4,
                         7800.0, 1, 2)
    O.or_apply_arc_pressure(P(o["Fz"]), P(o["flags"]), P(o["phi"]), ai, aj, ak, 6.0, 500.0, 2e-5,
                            2e-4, 7800.0, *dims)
    K.add_hydrostatic_gravity(g["Fz"], g["f_l"], g["flags"], 1.0, 3e-7, 1, 2)
    O.or_add_hydrostatic_gravity(P(o["Fz"]), P(o["f_l"]), P(o["flags"]), 1.0, 3e-7, *dims)
    K.add_buoyancy(g["T"], g["Fz"], g["f_l"], g["flags"], 3e-7, 1.2e-4, 1780.0, 1.0, 1, 2)
    O.or_add_buoyancy(P(o["T"]), P(o["Fz"]), P(o["f_l"]), P(o["flags"]), 3e-7, 1.2e-4, 1780.0,
                      1.0, *dims)
    K.apply_droplet_impact_pressure(g["Fz"], g["flags"], g["phi"], g["f_l"], ai, aj, ak, 3.0,
                                    2000.0, 2e-5, 2e-4, 7800.0, 1, 2)
    O.or_apply_droplet_impact_pressure(P(o["Fz"]), P(o["flags"]), P(o["phi"]), P(o["f_l"]), ai, aj,
                                       ak, 3.0, 2000.0, 2e-5, 2e-4, 7800.0, *dims)
    K.clamp_body_force_magnitude(g["Fx"], g["Fy"], g["Fz"], g["flags"], 0.02, 1, 2)
    O.or_clamp_body_force_magnitude(P(o["Fx"]), P(o["Fy"]), P(o["Fz"]), P(o["flags"]), 0.02, *dims)
    for k in ("Fx", "Fy", "Fz"):
        check(k, down(g[k]), o[k])


def test_curvature_and_velocity_clamp(env):
    dims = DIMS[0]
    s = random_state(*dims, seed=19)
    g = {k: up(env, s[k]) for k in ("phi", "flags", "ux", "uy", "uz")}
    o = {k: s[k].copy() for k in g}
    kap_g = env["torch"].zeros_like(g["phi"])
    kap_o = np.zeros_like(o["phi"])
    env["K"].compute_curvature_field(g["phi"], g["flags"], kap_g, 2, *dims)
    env["O"].or_compute_curvature_field(P(o["phi"]), P(o["flags"]), P(kap_o), *dims)
    check("kappa", down(kap_g), kap_o)
    env["K"].clamp_velocity_mach(g["ux"], g["uy"], g["uz"], g["flags"], 0.04, 1, 2)
    env["O"].or_clamp_velocity_mach(P(o["ux"]), P(o["uy"]), P(o["uz"]), P(o["flags"]), 0.04, *dims)
    for k in ("ux", "uy", "uz"):
        check(k, down(g[k]), o[k])


def test_freeze_kernels(env):
    dims = DIMS[0]
    s = random_state(*dims, seed=20)
    s["T"] = np.random.default_rng(2).uniform(1500, 2000, dims).astype(np.float32)
    keys = ("T", "H", "f_l", "phi", "flags", "ux", "uy", "uz")
    g = {k: up(env, s[k]) for k in keys}
    o = {k: s[k].copy() for k in keys}
    K, O = env["K"], env["O"]
    cp, Ts, Tl, L = 5.3e6, 1723.0, 1780.0, 2.1e9
    K.remelt_hot_solid_scalar(g["T"], g["H"], g["f_l"], g["phi"], g["flags"], cp, Ts, Tl, L, 1, 0)
    O.or_remelt_hot_solid_scalar(P(o["T"]), P(o["H"]), P(o["f_l"]), P(o["phi"]), P(o["flags"]), cp,
                                 Ts, Tl, L, *dims)
    K.solidify_trailing_pool_scalar(g["T"], g["H"], g["f_l"], g["phi"], g["flags"], g["ux"],
                                    g["uy"], g["uz"], cp, 14.0, 10.0, 4.0, 1.0, 0.0, 0.0, 3.0,
                                    1815.0, Ts, 1, 0, 2)
    O.or_solidify_trailing_pool_scalar(P(o["T"]), P(o["H"]), P(o["f_l"]), P(o["phi"]),
                                       P(o["flags"]), P(o["ux"]), P(o["uy"]), P(o["uz"]), cp, 14.0,
                                       10.0, 4.0, 1.0, 0.0, 0.0, 3.0, 1815.0, Ts, *dims)
    K.solidify_cooled_metal(g["T"], g["f_l"], g["phi"], g["flags"], g["ux"], g["uy"], g["uz"], Ts,
                            True, 1, 0, 2)
    O.or_solidify_cooled_metal(P(o["T"]), P(o["f_l"]), P(o["phi"]), P(o["flags"]), P(o["ux"]),
                               P(o["uy"]), P(o["uz"]), Ts, 1, *dims)
    for k in keys:
        check(k, down(g[k]), o[k])


@pytest.mark.parametrize("footprint", [4.0, 9.5])
def test_feed_wire_surface_single_call(env, footprint):
    """One reference feed_wire_surface call: serial-sweep semantics on the GPU wavefront."""
    dims = (40, 36, 30)
    s = random_state(*dims, seed=21, nz_solid=10)
    keys = ("f", "flags", "f_l", "phi", "H", "T", "rho")
    g = {k: up(env, s[k]) for k in keys}
    o = {k: s[k].copy() for k in keys}
    torch = env["torch"]
    ai, aj, ak = 19.3, 17.6, 10.0
    cell_vol = 8e-12
    for target in (40 * cell_vol, 400 * cell_vol):
        vg = torch.zeros((), dtype=torch.float32, device=env["dev"])
        rg = torch.zeros((), dtype=torch.float32, device=env["dev"])
        vo = np.zeros(1, np.float32)
        ro = np.zeros(1, np.float32)
        env["K"].feed_wire_surface(g["f"], g["flags"], g["f_l"], g["phi"], g["H"], g["T"], g["rho"],
                                   ai, aj, ak, footprint, 2.5, target, 2280.0, 5.3e6, 2.1e9, 1.0, vg,
                                   rg, cell_vol, 2, 0, 1, *dims)
        env["O"].or_feed_wire_surface(P(o["f"]), P(o["flags"]), P(o["f_l"]), P(o["phi"]), P(o["H"]),
                                      P(o["T"]), P(o["rho"]), ai, aj, ak, footprint, 2.5, target,
                                      2280.0, 5.3e6, 2.1e9, 1.0, P(vo), P(ro), cell_vol, *dims)
        assert float(vg.item()) == float(vo[0]) and float(rg.item()) == float(ro[0])
        for k in keys:
            check(k, down(g[k]), o[k])


def test_telemetry_reduce(env):
    dims = DIMS[1]
    s = random_state(*dims, seed=22)
    rng = np.random.default_rng(3)
    dT = (rng.standard_normal(dims) * 1e5).astype(np.float32)
    g = {k: up(env, s[k]) for k in ("T", "f_l", "phi", "ux", "uy", "uz")}
    torch = env["torch"]
    out_g = torch.zeros(8, dtype=torch.float64, device=env["dev"])
    sc = torch.zeros(__import__("waam_solver_amd._lib", fromlist=["lib"]).lib().scratch_bytes(),
                     dtype=torch.uint8, device=env["dev"])
    env["K"].telemetry_pool_reduce_into(g["T"], g["f_l"], g["phi"], up(env, dT), g["ux"], g["uy"],
                                        g["uz"], out_g, sc)
    out_o = np.zeros(8, np.float64)
    o = {k: s[k].copy() for k in g}
    env["O"].or_telemetry_pool_reduce(P(o["T"]), P(o["f_l"]), P(o["phi"]), P(dT), P(o["ux"]),
                                      P(o["uy"]), P(o["uz"]), P(out_o), *dims)
    assert np.array_equal(down(out_g), out_o), (down(out_g), out_o)


# Knot tables for the table-path kernels: three-knot cp/k/mu, two-knot dgamma/dT, and an
# eight-knot cp table with a zero-width interval (the lookup's last matching interval wins).
_TABLES = dict(cp=[(300.0, 650.0), (1000.0, 680.0), (1793.0, 720.0)],
               k=[(300.0, 28.0), (1000.0, 30.0), (1793.0, 34.0)],
               mu=[(300.0, 0.0075), (1793.0, 0.0065), (2000.0, 0.0055)],
               dgamma_dT=[(1793.0, -4.3e-4), (2000.0, -3.8e-4)])
_CP8 = [(300.0, 500.0), (600.0, 560.0), (900.0, 600.0), (900.0, 640.0), (1400.0, 690.0),
        (1700.0, 720.0), (2100.0, 760.0), (2600.0, 800.0)]


def _packed(tables):
    from oracle.oracle import Material, pack_tables

    m = Material(name="t", k=30.0, rho=7800.0, cp=680.0, T_solidus=1723.0, T_liquidus=1780.0,
                 L_fusion=2.7e5, gamma_0=1.8, dgamma_dT=-4.0e-4, mu=0.0065, tables=tables)
    return m, pack_tables(m)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("cp_table", ["three", "eight"])
@pytest.mark.parametrize("use_tables", [1, 0])
def test_table_kernels(env, dims, cp_table, use_tables):
    """Every temperature-dependent-material entry point on one seeded state, in step order."""
    s = random_state(*dims, seed=21)
    rng = np.random.default_rng(5)
    tables = dict(_TABLES, cp=_CP8) if cp_table == "eight" else dict(_TABLES)
    m, (knots, counts) = _packed(tables)
    # exact knot temperatures exercise the end clamps and the zero-width interval
    T = s["T"]
    hit = rng.random(dims) < 0.05
    T[hit] = rng.choice(np.array([t for t, _ in tables["cp"]] + [250.0, 2900.0], np.float32),
                        hit.sum())
    s["H"] = (T * np.float32(5.2e6)).astype(np.float32)
    keys = ("T", "H", "f_l", "phi", "flags", "ux", "uy", "uz", "Fx", "Fy", "Fz", "f", "f2",
            "rho")
    g = {k: up(env, s[k]) for k in keys}
    o = {k: s[k].copy() for k in keys}
    props = ("cp_rho", "alpha", "dgamma", "tau")
    for p in props:
        g[p] = up(env, np.zeros(dims, np.float32))
        o[p] = np.zeros(dims, np.float32)
    K, O = env["K"], env["O"]
    rho, dt, dx = 7800.0, 2e-5, 2e-4
    fscale = dt * dt / (rho * dx * dx) / dx
    refs = (5.3e6, 0.0471, -3.1e-7, 0.5125, 0.9)
    t = knots
    K.refresh_thermal_properties(
        g["T"], g["cp_rho"], g["alpha"], g["dgamma"], g["tau"], t[0:8], t[8:16], t[16:24],
        t[24:32], t[32:40], t[40:48], t[48:56], t[56:64], *[int(c) for c in counts], rho, dt,
        dx, m.cp, m.k, m.mu, m.dgamma_dT, fscale, *refs, use_tables, g["flags"], 1, 2)
    consts = np.array([rho, dt, dx, m.cp, m.k, m.mu, m.dgamma_dT, fscale, *refs], np.float32)
    O.or_refresh_thermal_properties(P(o["T"]), P(o["cp_rho"]), P(o["alpha"]), P(o["dgamma"]),
                                    P(o["tau"]), P(knots), P(counts), P(consts), use_tables,
                                    P(o["flags"]), *dims)
    for p in props:
        check(p, down(g[p]), o[p])
    K.advect_diffuse_temperature_variable(g["H"], g["T"], g["ux"], g["uy"], g["uz"], g["flags"],
                                          g["alpha"], g["cp_rho"], 1.0, 1, 2, *dims)
    H_in = o["H"].copy()
    O.or_advect_diffuse_temperature_variable_jacobi(
        P(H_in), P(o["H"]), P(o["T"]), P(o["ux"]), P(o["uy"]), P(o["uz"]), P(o["flags"]),
        P(o["alpha"]), P(o["cp_rho"]), 1.0, *dims)
    check("H advect", down(g["H"]), o["H"])
    K.clamp_enthalpy_floor(g["H"], g["cp_rho"], g["flags"], 1200.0, 2)
    O.or_clamp_enthalpy_floor(P(o["H"]), P(o["cp_rho"]), P(o["flags"]), 1200.0, *dims)
    check("H floor", down(g["H"]), o["H"])
    L_rho, Ts, Tl = 2.1e9, 1723.0, 1780.0
    H_sol = 7800.0 * 700.0 * Ts
    H_liq = H_sol + L_rho
    K.update_phase_variable_cp(g["H"], g["T"], g["f_l"], g["cp_rho"], L_rho, Ts, Tl, H_sol, H_liq)
    O.or_update_phase_variable_cp(P(o["H"]), P(o["T"]), P(o["f_l"]), P(o["cp_rho"]), L_rho, Ts,
                                  Tl, H_sol, H_liq, *dims)
    check("T phase", down(g["T"]), o["T"])
    check("f_l phase", down(g["f_l"]), o["f_l"])
    K.clamp_enthalpy_ceiling_variable_cp(g["H"], g["flags"], g["cp_rho"], H_liq, Tl, 2500.0, 2)
    O.or_clamp_enthalpy_ceiling_variable_cp(P(o["H"]), P(o["flags"]), P(o["cp_rho"]), H_liq, Tl,
                                            2500.0, *dims)
    check("H ceiling", down(g["H"]), o["H"])
    K.compute_marangoni_force_variable(g["T"], g["phi"], g["f_l"], g["Fx"], g["Fy"], g["Fz"],
                                       g["flags"], g["dgamma"], 1, 2, *dims)
    O.or_compute_marangoni_force_variable(P(o["T"]), P(o["phi"]), P(o["f_l"]), P(o["Fx"]),
                                          P(o["Fy"]), P(o["Fz"]), P(o["flags"]), P(o["dgamma"]),
                                          *dims)
    for k in ("Fx", "Fy", "Fz"):
        check(k, down(g[k]), o[k])
    K.collide_srt_variable_tau(g["f"], g["f2"], g["rho"], g["ux"], g["uy"], g["uz"], g["Fx"],
                               g["Fy"], g["Fz"], g["f_l"], g["flags"], g["tau"], 1.6e5, 1, 2,
                               *dims)
    O.or_collide_srt_variable_tau(P(o["f"]), P(o["f2"]), P(o["rho"]), P(o["ux"]), P(o["uy"]),
                                  P(o["uz"]), P(o["Fx"]), P(o["Fy"]), P(o["Fz"]), P(o["f_l"]),
                                  P(o["flags"]), P(o["tau"]), 1.6e5, *dims)
    for k in ("f2", "rho", "ux", "uy", "uz"):
        check(k, down(g[k]), o[k])
    K.remelt_hot_solid(g["T"], g["H"], g["f_l"], g["phi"], g["flags"], g["cp_rho"], Ts, Tl,
                       L_rho, 1, 0)
    O.or_remelt_hot_solid(P(o["T"]), P(o["H"]), P(o["f_l"]), P(o["phi"]), P(o["flags"]),
                          P(o["cp_rho"]), Ts, Tl, L_rho, *dims)
    K.solidify_trailing_pool(g["T"], g["H"], g["f_l"], g["phi"], g["flags"], g["ux"], g["uy"],
                             g["uz"], g["cp_rho"], dims[0] / 2, dims[1] / 2, 4.0, 1.0, 0.0, 0.0,
                             3.0, 2400.0, Ts, 1, 0, 2)
    O.or_solidify_trailing_pool(P(o["T"]), P(o["H"]), P(o["f_l"]), P(o["phi"]), P(o["flags"]),
                                P(o["ux"]), P(o["uy"]), P(o["uz"]), P(o["cp_rho"]), dims[0] / 2,
                                dims[1] / 2, 4.0, 1.0, 0.0, 0.0, 3.0, 2400.0, Ts, *dims)
    for k in ("T", "H", "f_l", "phi", "flags", "ux", "uy", "uz"):
        check(k, down(g[k]), o[k])


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("n_shift", [5, 8, 80])
def test_shift_window(env, dims, n_shift):
    """shift_simulation_window_x on seeded fields, n_shift below and above nx."""
    import ctypes

    from oracle.oracle import OracleTwin

    rng = np.random.default_rng(n_shift)
    names = OracleTwin.WINDOW_FIELDS
    o = {n: rng.standard_normal(dims).astype(np.float32) for n in names}
    o["flags"] = rng.integers(0, 5, dims).astype(np.uint8)
    o["f_a"] = rng.standard_normal((19,) + dims).astype(np.float32)
    o["f_b"] = rng.standard_normal((19,) + dims).astype(np.float32)
    g = {k: up(env, v) for k, v in o.items()}
    env["K"].shift_simulation_window_x(n_shift, g["f_a"], g["f_b"], *(g[n] for n in names[:4]),
                                       g["flags"], *(g[n] for n in names[4:]), 293.0, 5.3e6, 1.0,
                                       3, 1, 2, 0, *dims)
    table = (ctypes.c_void_p * len(names))(*[o[n].ctypes.data for n in names])
    env["O"].or_shift_simulation_window_x(n_shift, P(o["f_a"]), P(o["f_b"]), table, P(o["flags"]),
                                          293.0, 5.3e6, 1.0, 3, *dims)
    for k in o:
        check(k, down(g[k]), o[k])


@pytest.mark.parametrize("dims", DIMS)
def test_full_tier_surface_forces(env, dims):
    """Vapour recoil (both forms) and gas shear on a seeded state with T spread across the
    boiling point."""
    s = random_state(*dims, seed=31)
    s["T"] = np.random.default_rng(3).uniform(2500.0, 4200.0, dims).astype(np.float32)
    keys = ("Fx", "Fy", "Fz", "T", "phi", "flags")
    g = {k: up(env, s[k]) for k in keys}
    o = {k: s[k].copy() for k in keys}
    K, O = env["K"], env["O"]
    ai, aj, ak, sig = dims[0] / 2 + 0.3, dims[1] / 2 - 0.4, 6.0, 3.5
    K.apply_vapor_recoil(g["Fz"], g["T"], g["phi"], g["flags"], ai, aj, ak, sig, 5000.0, 1780.0,
                         2e-5, 2e-4, 7800.0, 1, 2)
    O.or_apply_vapor_recoil(P(o["Fz"]), P(o["T"]), P(o["phi"]), P(o["flags"]), ai, aj, sig, 5000.0,
                            1780.0, 2e-5, 2e-4, 7800.0, *dims)
    check("Fz recoil", down(g["Fz"]), o["Fz"])
    K.apply_vapor_recoil_clausius_clapeyron(g["Fz"], g["T"], g["phi"], g["flags"], ai, aj, ak, sig,
                                            101325.0, 3100.0, 6e6, 450.0, 0.54, 2e-5, 2e-4,
                                            7800.0, 1, 2)
    O.or_apply_vapor_recoil_cc(P(o["Fz"]), P(o["T"]), P(o["phi"]), P(o["flags"]), ai, aj, sig,
                               101325.0, 3100.0, 6e6, 450.0, 0.54, 2e-5, 2e-4, 7800.0, *dims)
    check("Fz recoil CC", down(g["Fz"]), o["Fz"])
    K.apply_gas_shear_stress(g["Fx"], g["Fy"], g["phi"], g["flags"], ai, aj, ak, sig, 115.2, 2e-5,
                             2e-4, 7800.0, 1, 2)
    O.or_apply_gas_shear_stress(P(o["Fx"]), P(o["Fy"]), P(o["phi"]), P(o["flags"]), ai, aj, sig,
                                115.2, 2e-5, 2e-4, 7800.0, *dims)
    check("Fx shear", down(g["Fx"]), o["Fx"])
    check("Fy shear", down(g["Fy"]), o["Fy"])


@pytest.mark.parametrize("dims", DIMS)
def test_lorentz_kernels(env, dims):
    """Every Lorentz entry on seeded fields, in solve order."""
    s = random_state(*dims, seed=41)
    rng = np.random.default_rng(41)
    K, O = env["K"], env["O"]
    n = dims
    g = {k: up(env, s[k]) for k in ("f_l", "flags", "Fx", "Fy", "Fz")}
    o = {k: s[k].copy() for k in g}
    z = lambda: np.zeros(n, np.float32)  # noqa: E731
    for name in ("sigma", "src", "Jx", "Jy", "Jz", "Bx", "By", "Bz"):
        o[name] = z()
        g[name] = up(env, o[name])
    o["phi"] = (rng.standard_normal(n) * 1e-3).astype(np.float32)
    g["phi"] = up(env, o["phi"])
    o["phi2"] = z()
    g["phi2"] = up(env, o["phi2"])
    K.elec_build_sigma(g["sigma"], g["f_l"], g["flags"], 1e6, 1e5, 2)
    O.or_elec_build_sigma(P(o["sigma"]), P(o["f_l"]), P(o["flags"]), 1e6, 1e5, *n)
    check("sigma", down(g["sigma"]), o["sigma"])
    ai, aj, ak, sig, dx = n[0] / 2 + 0.3, n[1] / 2 - 0.2, 5.5, 2.5, 2e-4
    K.elec_clear_source(g["src"])
    K.elec_inject_arc_source(g["src"], ai, aj, ak, sig, 140.0, dx, n[2])
    O.or_elec_inject_arc_source(P(o["src"]), ai, aj, ak, sig, 140.0, dx, *n)
    check("source", down(g["src"]), o["src"])
    K.elec_normalize_source(g["src"], 140.0, dx)
    O.or_elec_normalize_source(P(o["src"]), 140.0, dx, *n)
    check("source normalised", down(g["src"]), o["src"])
    K.elec_jacobi_step(g["phi"], g["phi2"], g["sigma"], g["src"], g["flags"], dx, 0.85, 2, 1, 4,
                       *n)
    O.or_elec_jacobi_step(P(o["phi"]), P(o["phi2"]), P(o["sigma"]), P(o["src"]), P(o["flags"]),
                          dx, 0.85, 4, *n)
    check("phi jacobi", down(g["phi2"]), o["phi2"])
    out_g = up(env, np.zeros(2, np.float32))
    out_o = np.zeros(2, np.float32)
    K.elec_l1_diff_into(g["phi2"], g["phi"], out_g)
    O.or_elec_l1_diff(P(o["phi2"]), P(o["phi"]), P(out_o), *n)
    check("l1", down(out_g), out_o)
    K.elec_compute_J(g["Jx"], g["Jy"], g["Jz"], g["phi2"], g["sigma"], g["flags"], dx, 2, *n)
    O.or_elec_compute_J(P(o["Jx"]), P(o["Jy"]), P(o["Jz"]), P(o["phi2"]), P(o["sigma"]),
                        P(o["flags"]), dx, *n)
    for k in ("Jx", "Jy", "Jz"):
        check(k, down(g[k]), o[k])
    nb, dr = 64, max(max(n[0], n[1]) / 128.0, 0.5)
    bins_o = np.zeros((n[2], nb), np.float32)
    bins_g = up(env, bins_o)
    K.elec_bin_axial_current(g["Jz"], bins_g, ai, aj, dx, dr, nb, n[2])
    O.or_elec_bin_axial_current(P(o["Jz"]), P(bins_o), ai, aj, dx, dr, nb, *n)
    check("bins", down(bins_g), bins_o)
    K.elec_prefix_bins(bins_g, nb, n[2])
    O.or_elec_prefix_bins(P(bins_o), nb, n[2])
    check("bins prefix", down(bins_g), bins_o)
    mu0 = 4e-7 * math.pi
    K.elec_B_axisymmetric(g["Bx"], g["By"], g["Bz"], bins_g, ai, aj, dx, dr, mu0, nb, n[2])
    O.or_elec_B_axisymmetric(P(o["Bx"]), P(o["By"]), P(o["Bz"]), P(bins_o), ai, aj, dx, dr, mu0,
                             nb, *n)
    for k in ("Bx", "By", "Bz"):
        check(k, down(g[k]), o[k])
    K.apply_lorentz_JxB(g["Fx"], g["Fy"], g["Fz"], g["Jx"], g["Jy"], g["Jz"], g["Bx"], g["By"],
                        g["Bz"], g["f_l"], g["flags"], 2e-5, dx, 7800.0, 2)
    O.or_apply_lorentz_JxB(*(P(o[k]) for k in ("Fx", "Fy", "Fz", "Jx", "Jy", "Jz", "Bx", "By",
                                                 "Bz", "f_l", "flags")), 2e-5, dx, 7800.0, *n)
    for k in ("Fx", "Fy", "Fz"):
        check(k, down(g[k]), o[k])
