"""Pinning the oracle's temperature-dependent material path.

Two kinds of checks, both independent of the restated code:
  * the per-cell property refresh against a float64 evaluation of the definitions (numpy's
    piecewise-linear interpolation for the knot lookup; alpha = k/(rho cp) dt/dx^2,
    tau = 3 (mu/rho) dt/dx^2 + 1/2, dgamma_lu = dgamma/dT * scale * marangoni_scale) and literal
    values at the lookup's special points (end clamps, a zero-width interval, an empty table);
  * every per-cell-cp kernel with a uniform cp_rho field against the scalar kernel of the same
    reference step, which the scalar pinning tests already fix: the table path must reduce to
    it bit for bit when the tables are flat.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import FLUID, GAS, IFACE, SOLID, Material, interp_knots, lib, pack_tables, ptr

RHO, DT, DX = 7800.0, 2e-5, 2e-4
REFS = dict(force_scale=DT * DT / (RHO * DX * DX) / DX, cp_rho_ref=5.3e6, alpha_ref=0.0471,
            dgamma_ref=-3.1e-7, tau_ref=0.5125, mscale=0.9)
TABLES = dict(cp=[(300.0, 650.0), (1000.0, 680.0), (1793.0, 720.0)],
              k=[(300.0, 28.0), (1000.0, 30.0), (1793.0, 34.0)],
              mu=[(300.0, 0.0075), (1793.0, 0.0065), (2000.0, 0.0055)],
              dgamma_dT=[(1793.0, -4.3e-4), (2000.0, -3.8e-4)])


def _material(tables):
    return Material(name="t", k=30.0, rho=RHO, cp=680.0, T_solidus=1723.0, T_liquidus=1780.0,
                    L_fusion=2.7e5, gamma_0=1.8, dgamma_dT=-4.0e-4, mu=0.0065, tables=tables)


def _refresh(T, flags, tables, use_tables=1):
    m = _material(tables)
    knots, counts = pack_tables(m)
    consts = np.array([RHO, DT, DX, m.cp, m.k, m.mu, m.dgamma_dT, REFS["force_scale"],
                       REFS["cp_rho_ref"], REFS["alpha_ref"], REFS["dgamma_ref"], REFS["tau_ref"],
                       REFS["mscale"]], np.float32)
    out = [np.zeros(T.shape, np.float32) for _ in range(4)]
    lib().or_refresh_thermal_properties(ptr(T), *(ptr(a) for a in out), ptr(knots), ptr(counts),
                                        ptr(consts), use_tables, ptr(flags), T.size, 1, 1)
    return m, out


def _interp(T, knots):
    t = np.array([p[0] for p in knots])
    v = np.array([p[1] for p in knots])
    return np.interp(T.astype(np.float64), t, v)


def test_refresh_matches_float64_definitions():
    rng = np.random.default_rng(3)
    T = rng.uniform(150.0, 2400.0, 4000).astype(np.float32)
    flags = np.full(T.size, FLUID, np.uint8)
    flags[::7] = IFACE
    m, (cp_rho, alpha, dgamma, tau) = _refresh(T, flags, TABLES)
    cp = _interp(T, TABLES["cp"])
    k = _interp(T, TABLES["k"])
    mu = _interp(T, TABLES["mu"])
    dg = _interp(T, TABLES["dgamma_dT"])
    np.testing.assert_allclose(cp_rho, RHO * cp, rtol=2e-7)
    np.testing.assert_allclose(alpha, k / (RHO * cp) * DT / DX ** 2, rtol=1e-6)
    np.testing.assert_allclose(tau, 3.0 * mu / RHO * DT / DX ** 2 + 0.5, rtol=1e-6)
    np.testing.assert_allclose(dgamma, dg * REFS["force_scale"] * REFS["mscale"], rtol=1e-6)


def test_refresh_solid_gas_and_disabled_tables_take_the_references():
    T = np.array([500.0, 1500.0, 2500.0, 900.0], np.float32)
    flags = np.array([SOLID, GAS, FLUID, IFACE], np.uint8)
    _, (cp_rho, alpha, dgamma, tau) = _refresh(T, flags, TABLES)
    f32 = np.float32
    for a, ref in ((cp_rho, "cp_rho_ref"), (alpha, "alpha_ref"), (dgamma, "dgamma_ref"),
                   (tau, "tau_ref")):
        assert a[0] == f32(REFS[ref]) and a[1] == f32(REFS[ref])
    _, (cp_rho, alpha, dgamma, tau) = _refresh(T, flags, TABLES, use_tables=0)
    assert np.all(cp_rho == f32(REFS["cp_rho_ref"])) and np.all(tau == f32(REFS["tau_ref"]))
    # tables off: the scalar dgamma still carries marangoni_scale on metal cells only
    assert dgamma[2] == f32(REFS["dgamma_ref"]) * f32(REFS["mscale"])
    assert dgamma[0] == f32(REFS["dgamma_ref"])


def test_lookup_special_points():
    """End clamps, a zero-width interval (the last matching interval wins, so the knot takes
    the value on its right), a single knot and an empty table (the packed (300 K, scalar))."""
    cp8 = [(300.0, 500.0), (600.0, 560.0), (900.0, 600.0), (900.0, 640.0), (1400.0, 690.0)]
    T = np.array([100.0, 300.0, 600.0, 900.0, 1400.0, 5000.0], np.float32)
    flags = np.full(T.size, FLUID, np.uint8)
    _, (cp_rho, *_r) = _refresh(T, flags, dict(TABLES, cp=cp8))
    expect = np.float32(RHO) * np.array([500.0, 500.0, 560.0, 640.0, 690.0, 690.0], np.float32)
    np.testing.assert_array_equal(cp_rho, expect)
    # the host lookup (used for H_sol) takes the first matching interval instead
    assert interp_knots(900.0, cp8, 0.0) == 600.0
    _, (cp_rho, *_r) = _refresh(T, flags, dict(TABLES, cp=[(1000.0, 700.0)]))
    assert np.all(cp_rho == np.float32(RHO) * np.float32(700.0))
    _, (cp_rho, alpha, dgamma, tau) = _refresh(T, flags, dict(TABLES, cp=[]))
    assert np.all(cp_rho == np.float32(RHO) * np.float32(680.0))


def test_pack_tables_layout():
    m = _material(dict(TABLES, k=[]))
    knots, counts = pack_tables(m)
    assert list(counts) == [3, 1, 3, 2]
    assert knots[0] == 300.0 and knots[8] == 650.0 and knots[2] == 1793.0 and knots[10] == 720.0
    assert knots[16] == 300.0 and knots[24] == np.float32(30.0)       # empty k -> (300, k)
    assert knots[48] == 1793.0 and knots[57] == np.float32(-3.8e-4)
    over = [(300.0 + 10 * i, float(i)) for i in range(11)]
    _, counts = pack_tables(_material(dict(TABLES, cp=over)))
    assert counts[0] == 8                                            # at most eight knots


def _state(n=(12, 10, 9), seed=4):
    rng = np.random.default_rng(seed)
    flags = rng.choice(np.array([FLUID, SOLID, GAS, IFACE], np.uint8), n, p=[.4, .3, .15, .15])
    T = rng.uniform(300.0, 3000.0, n).astype(np.float32)
    H = (T * np.float32(5.3e6) * rng.uniform(0.7, 1.3, n)).astype(np.float32)
    phi = rng.uniform(0, 1, n).astype(np.float32)
    f_l = rng.uniform(0, 1, n).astype(np.float32)
    u = [(rng.standard_normal(n) * 0.03).astype(np.float32) for _ in range(3)]
    return dict(flags=flags, T=T, H=H, phi=phi, f_l=f_l, ux=u[0], uy=u[1], uz=u[2])


@pytest.mark.parametrize("seed", [4, 5])
def test_uniform_cp_reduces_to_scalar_path(seed):
    L = lib()
    s = _state(seed=seed)
    dims = s["T"].shape
    cp = np.float32(5.3e6)
    cpf = np.full(dims, cp, np.float32)
    f32 = np.float32
    Ts, Tl, Lr = f32(1723.0), f32(1780.0), f32(2.1e9)
    H_sol = cp * Ts
    H_liq = f32(H_sol + Lr)

    def two():
        return {k: v.copy() for k, v in s.items()}, {k: v.copy() for k, v in s.items()}

    a, b = two()
    alpha = np.full(dims, np.float32(0.05), np.float32)
    Ha, Hb = a["H"].copy(), b["H"].copy()
    L.or_advect_diffuse_temperature_variable_jacobi(
        ptr(a["H"]), ptr(Ha), ptr(a["T"]), ptr(a["ux"]), ptr(a["uy"]), ptr(a["uz"]),
        ptr(a["flags"]), ptr(alpha), ptr(cpf), 1.0, *dims)
    L.or_advect_diffuse_temperature_jacobi(ptr(b["H"]), ptr(Hb), ptr(b["T"]), ptr(b["ux"]),
                                           ptr(b["uy"]), ptr(b["uz"]), ptr(b["flags"]), 0.05, 1.0,
                                           float(cp), *dims)
    np.testing.assert_array_equal(Ha, Hb)

    a, b = two()
    L.or_clamp_enthalpy_floor(ptr(a["H"]), ptr(cpf), ptr(a["flags"]), 1200.0, *dims)
    L.or_clamp_enthalpy_floor_scalar(ptr(b["H"]), float(cp), ptr(b["flags"]), 1200.0, *dims)
    np.testing.assert_array_equal(a["H"], b["H"])

    a, b = two()
    L.or_update_phase_variable_cp(ptr(a["H"]), ptr(a["T"]), ptr(a["f_l"]), ptr(cpf), float(Lr),
                                  float(Ts), float(Tl), float(H_sol), float(H_liq), *dims)
    L.or_update_phase(ptr(b["H"]), ptr(b["T"]), ptr(b["f_l"]), float(cp), float(Lr), float(Ts),
                      float(Tl), *dims)
    np.testing.assert_array_equal(a["T"], b["T"])
    np.testing.assert_array_equal(a["f_l"], b["f_l"])

    a, b = two()
    L.or_clamp_enthalpy_ceiling_variable_cp(ptr(a["H"]), ptr(a["flags"]), ptr(cpf), float(H_liq),
                                            float(Tl), 2500.0, *dims)
    L.or_clamp_enthalpy_ceiling_scalar(ptr(b["H"]), ptr(b["flags"]), float(cp), float(Ts),
                                       float(Tl), 2500.0, float(Lr), *dims)
    np.testing.assert_array_equal(a["H"], b["H"])

    a, b = two()
    dgf = np.full(dims, np.float32(-3.1e-7), np.float32)
    Fa = [np.zeros(dims, np.float32) for _ in range(3)]
    Fb = [np.zeros(dims, np.float32) for _ in range(3)]
    L.or_compute_marangoni_force_variable(ptr(a["T"]), ptr(a["phi"]), ptr(a["f_l"]),
                                          *(ptr(x) for x in Fa), ptr(a["flags"]), ptr(dgf), *dims)
    L.or_compute_marangoni_force(ptr(b["T"]), ptr(b["phi"]), ptr(b["f_l"]), *(ptr(x) for x in Fb),
                                 ptr(b["flags"]), float(np.float32(-3.1e-7)), *dims)
    for x, y in zip(Fa, Fb):
        np.testing.assert_array_equal(x, y)
    assert any(np.any(x != 0) for x in Fa)

    a, b = two()
    L.or_remelt_hot_solid(ptr(a["T"]), ptr(a["H"]), ptr(a["f_l"]), ptr(a["phi"]), ptr(a["flags"]),
                          ptr(cpf), float(Ts), float(Tl), float(Lr), *dims)
    L.or_remelt_hot_solid_scalar(ptr(b["T"]), ptr(b["H"]), ptr(b["f_l"]), ptr(b["phi"]),
                                 ptr(b["flags"]), float(cp), float(Ts), float(Tl), float(Lr), *dims)
    L.or_solidify_trailing_pool(ptr(a["T"]), ptr(a["H"]), ptr(a["f_l"]), ptr(a["phi"]),
                                ptr(a["flags"]), ptr(a["ux"]), ptr(a["uy"]), ptr(a["uz"]), ptr(cpf),
                                6.0, 5.0, 3.0, 1.0, 0.0, 0.0, 2.0, 2400.0, float(Ts), *dims)
    L.or_solidify_trailing_pool_scalar(ptr(b["T"]), ptr(b["H"]), ptr(b["f_l"]), ptr(b["phi"]),
                                       ptr(b["flags"]), ptr(b["ux"]), ptr(b["uy"]), ptr(b["uz"]),
                                       float(cp), 6.0, 5.0, 3.0, 1.0, 0.0, 0.0, 2.0, 2400.0,
                                       float(Ts), *dims)
    for k in s:
        np.testing.assert_array_equal(a[k], b[k])
    assert np.any(a["flags"] != s["flags"])


def test_table_twin_first_welding_step_clamps_against_zero_cp():
    """The property fields are zero until the first refresh: a first step that welds clamps the
    enthalpy of every non-gas cell below H_liq to cp_rho_field * T_cap = 0 (reference
    solvers/coupled_step.py:104-108 runs before the refresh at :168-182)."""
    from oracle.oracle import OracleTwin

    ot = OracleTwin(_material(TABLES), nx=24, ny=16, nz=12, dx=2e-4, droplet_freq_hz=0.0,
                    max_tracers=16, enable_heat_loss=False)
    ot.reset()
    metal = ot.g.flags != GAS
    ot.step(12 * 2e-4, 8 * 2e-4, True)
    below = metal & (ot.g.T < 1780.0)
    assert below.sum() > 0 and np.all(ot.g.H[below] == 0.0) and np.all(ot.g.T[below] == 0.0)
    assert np.all(ot.g.cp_rho_field[metal & (ot.g.flags != SOLID)] > 0)
