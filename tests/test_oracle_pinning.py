"""Pinning the oracle: checks that fix its values from outside the restatement.

Each check compares the oracle's fp32 output with something derived independently of the
restated code: moment identities of the collision operators worked out from their published
definitions (SURVEY-level formulas in the reference's docstrings: cumulant_kernel.py:28-209,
kernels.py:2778-2885; Guo et al. 2002 for the forcing), brute-force definitions on tiny grids
(stream, telemetry), closed forms per cell (losses, arc pressure, buoyancy, clamps), and literal
expected values at threshold boundaries (flag classification, contact-angle rules, freeze
rules). A plausible mistake in the restatement (a dropped term, a wrong weight, a swapped
relaxation rate, a strict/non-strict comparison, a wrong index) fails one of them.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.oracle import EX, EY, EZ, FLUID, GAS, IFACE, OPP, SOLID, W, lib, ptr

E = np.array([EX, EY, EZ], dtype=np.float64)  # (3, 19)
Wd = np.array(W, dtype=np.float64)


def _cells(n, seed, rho_spread=0.02, u_max=0.05, neq=0.01):
    """n random near-equilibrium population sets (19, n) in fp32."""
    rng = np.random.default_rng(seed)
    rho = 1.0 + rho_spread * rng.uniform(-1, 1, n)
    u = u_max * rng.uniform(-1, 1, (3, n))
    eu = E.T @ u                                        # (19, n)
    feq = Wd[:, None] * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (u ** 2).sum(0))
    f = feq * (1 + neq * rng.uniform(-1, 1, feq.shape))
    return f.astype(np.float32)


def _collide(kind, f, F, fl, omega_s, omega_b=None, C=0.0):
    n = f.shape[1]
    L = lib()
    f_dst = np.zeros_like(f)
    rho, ux, uy, uz = (np.zeros(n, np.float32) for _ in range(4))
    Fx, Fy, Fz = (np.ascontiguousarray(F[a].astype(np.float32)) for a in range(3))
    flv = np.ascontiguousarray(fl.astype(np.float32))
    flags = np.full(n, FLUID, np.uint8)
    if kind == "mrt":
        L.or_collide_mrt(ptr(f), ptr(f_dst), ptr(rho), ptr(ux), ptr(uy), ptr(uz), ptr(Fx), ptr(Fy),
                         ptr(Fz), ptr(flv), ptr(flags), omega_s, omega_b, C, n, 1, 1)
    else:
        tau = 1.0 / omega_s
        L.or_collide_srt(ptr(f), ptr(f_dst), ptr(rho), ptr(ux), ptr(uy), ptr(uz), ptr(Fx), ptr(Fy),
                         ptr(Fz), ptr(flv), ptr(flags), tau, omega_s, 1.0, C, n, 1, 1)
    return f_dst, rho, np.stack([ux, uy, uz])


# ---------------------------------------------------------------------------- collision
@pytest.mark.parametrize("omega_s,omega_b", [(1.2, 1.2), (1.7, 0.9), (0.8, 1.5)])
def test_mrt_moments_follow_the_regularized_reconstruction(omega_s, omega_b):
    """Moments of f' against the closed forms of the reference's reconstruction
    f'_q = w_q[rho + 3 rho e.u + 4.5 rho (e.u)^2 - 1.5 rho u^2 + 4.5 e e : D]
           + w_q (1 - w_s/2)[3 e.F + 9 (e.F)(e.u) - 3 u.F],
    D = k' - rho/3 I, k = P - rho u u, trace of k relaxed at w_b, the rest at w_s. With the
    D3Q19 weight moments (sum w e e = I/3, sum w e_a e_b e_c e_d = (dd + dd + dd)/9):
      sum f'     = rho + 1.5 tr D
      sum e f'   = rho u + (1 - w_s/2) F
      sum e e f' = rho/3 I + rho u u + D + 0.5 tr(D) I + (1 - w_s/2)(u F + F u)."""
    n = 64
    rng = np.random.default_rng(5)
    f = _cells(n, 1)
    F = 1e-4 * rng.uniform(-1, 1, (3, n))
    fo, rho, u = _collide("mrt", f, F, np.ones(n), omega_s, omega_b)
    fd, fo = f.astype(np.float64), fo.astype(np.float64)
    r, u, F = rho.astype(np.float64), u.astype(np.float64), F.astype(np.float32).astype(np.float64)
    P = np.einsum("aq,bq,qn->abn", E, E, fd)
    k = P - r * np.einsum("an,bn->abn", u, u)
    km = np.trace(k) / 3.0
    kmp = km - omega_b * (km - r / 3.0)
    I = np.eye(3)[:, :, None]
    kp = kmp * I + (1 - omega_s) * (k - km * I)
    D = kp - (r / 3.0) * I
    trD = np.trace(D)
    gs = 1 - omega_s / 2
    m0 = fo.sum(0)
    m1 = E @ fo
    m2 = np.einsum("aq,bq,qn->abn", E, E, fo)
    exp_m2 = (r / 3.0) * I + r * np.einsum("an,bn->abn", u, u) + D + 0.5 * trD * I + gs * (
        np.einsum("an,bn->abn", u, F) + np.einsum("an,bn->abn", F, u))
    assert np.allclose(m0, r + 1.5 * trD, rtol=0, atol=2e-6)
    assert np.allclose(m1, r * u + gs * F, rtol=0, atol=2e-6)
    assert np.allclose(m2, exp_m2, rtol=0, atol=3e-6)
    assert np.allclose(r, fd.sum(0), rtol=0, atol=2e-6)


def test_mrt_rest_state_is_fixed():
    n = 8
    rho = np.linspace(0.98, 1.02, n)
    f = (Wd[:, None] * rho).astype(np.float32)
    fo, r, u = _collide("mrt", f, np.zeros((3, n)), np.ones(n), 1.3, 0.7)
    assert np.abs(u).max() == 0.0
    assert np.allclose(fo, f, rtol=1e-6, atol=0)   # fp32 rounding of the moment sums


@pytest.mark.parametrize("omega", [0.6, 1.0, 1.85])
def test_srt_is_bgk_with_guo_forcing(omega):
    """Guo, Zheng & Shi (2002): f' = f - w (f - f_eq(rho, u)) + (1 - w/2) w_q [3 (e - u).F
    + 9 (e.u)(e.F)], u = (j + F/2)/rho, written from the paper in fp64; also j' = j + F."""
    n = 64
    rng = np.random.default_rng(9)
    f = _cells(n, 2)
    F = (1e-4 * rng.uniform(-1, 1, (3, n))).astype(np.float32).astype(np.float64)
    fo, rho, u_out = _collide("srt", f, F, np.ones(n), omega)
    fd = f.astype(np.float64)
    r = fd.sum(0)
    j = E @ fd
    u = (j + 0.5 * F) / r
    eu = E.T @ u
    feq = Wd[:, None] * r * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (u ** 2).sum(0))
    eF = E.T @ F
    S = (1 - omega / 2) * Wd[:, None] * (3 * (eF - (u * F).sum(0)) + 9 * eu * eF)
    expect = fd - omega * (fd - feq) + S
    assert np.allclose(fo, expect, rtol=0, atol=3e-7)
    assert np.allclose(u_out, u, rtol=0, atol=1e-7)
    assert np.allclose(fo.astype(np.float64).sum(0), r, atol=2e-6)
    assert np.allclose(E @ fo.astype(np.float64), j + F, atol=2e-6)


def test_darcy_damping_scales_velocity():
    """u = u_raw / (1 + C (1 - fl)^2 / (fl^3 + 1e-6)) (kernels.py:2815-2820)."""
    n = 16
    f = _cells(n, 3)
    fl = np.linspace(0.05, 1.0, n)
    C = 1.6e5
    _, _, u = _collide("srt", f, np.zeros((3, n)), fl, 1.1, C=C)
    _, _, u0 = _collide("srt", f, np.zeros((3, n)), np.ones(n), 1.1, C=0.0)
    ck = 1.0 + C * (1 - fl) ** 2 / (fl ** 3 + 1e-6)
    assert np.allclose(u, u0 / ck, rtol=3e-6, atol=1e-12)


# ---------------------------------------------------------------------------- stream
def test_stream_matches_brute_force_definition():
    """Pull, x periodic, y/z clamped; GAS pulls unconditionally; other non-SOLID cells bounce
    back off SOLID neighbours; SOLID untouched (kernels.py:2971-3010), on a 5x4x3 grid."""
    nx, ny, nz = 5, 4, 3
    N = nx * ny * nz
    rng = np.random.default_rng(4)
    f = rng.random((19, nx, ny, nz)).astype(np.float32)
    flags = rng.choice(np.array([FLUID, SOLID, GAS, IFACE], np.uint8), (nx, ny, nz))
    out = np.full_like(f, -7.0)
    lib().or_stream(ptr(f), ptr(out), ptr(flags), nx, ny, nz)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                fl = flags[i, j, k]
                for q in range(19):
                    if fl == SOLID:
                        want = -7.0
                    else:
                        n_ = ((i - EX[q]) % nx, min(max(j - EY[q], 0), ny - 1),
                              min(max(k - EZ[q], 0), nz - 1))
                        if fl != GAS and flags[n_] == SOLID:
                            want = f[OPP[q], i, j, k]
                        else:
                            want = f[(q,) + n_]
                    assert out[q, i, j, k] == np.float32(want), (q, i, j, k)
    assert N == out[0].size


def test_stream_of_uniform_fluid_is_identity():
    nx, ny, nz = 6, 5, 4
    f = np.broadcast_to(np.arange(19, dtype=np.float32)[:, None, None, None],
                        (19, nx, ny, nz)).copy()
    flags = np.full((nx, ny, nz), FLUID, np.uint8)
    out = np.zeros_like(f)
    lib().or_stream(ptr(f), ptr(out), ptr(flags), nx, ny, nz)
    assert np.array_equal(out, f)


# ---------------------------------------------------------------------------- flags
# (phi, f_l, k < nz_solid) -> expected flag, at the thresholds of kernels.py:2406-2419:
# SOLID if k < nz_solid; SOLID if phi > 0.85 and f_l < 0.05; FLUID if phi > 0.55 and
# f_l > 0.1; GAS if phi < 0.05; else IFACE. Values written out by hand, not computed.
FLAG_TABLE = [
    (0.0, 0.0, True, SOLID), (0.5, 0.9, True, SOLID),
    (0.86, 0.04, False, SOLID), (0.85, 0.04, False, IFACE), (0.86, 0.05, False, IFACE),
    (0.86, 0.11, False, FLUID), (0.56, 0.11, False, FLUID), (0.55, 0.11, False, IFACE),
    (0.56, 0.10, False, IFACE), (0.049, 0.0, False, GAS), (0.05, 0.0, False, IFACE),
    (0.0, 0.9, False, GAS), (0.3, 0.3, False, IFACE), (1.0, 1.0, False, FLUID),
]


def test_flag_classification_truth_table():
    n = len(FLAG_TABLE)
    # one column per row: k = 0 is below nz_solid = 1, k = 1 is above
    phi = np.zeros((n, 1, 2), np.float32)
    fl = np.zeros((n, 1, 2), np.float32)
    for r, (p, f_, below, _) in enumerate(FLAG_TABLE):
        kk = 0 if below else 1
        phi[r, 0, kk] = p
        fl[r, 0, kk] = f_
    flags = np.zeros((n, 1, 2), np.uint8)
    lib().or_update_flags_from_phi(ptr(phi), ptr(fl), ptr(flags), 1, n, 1, 2)
    for r, (_, _, below, want) in enumerate(FLAG_TABLE):
        assert flags[r, 0, 0 if below else 1] == want, FLAG_TABLE[r]


def test_phi_liquid_fraction_sync_cases():
    """f_l := 0 where GAS, phi < 0.05 or SOLID (kernels.py:3484-3504)."""
    phi = np.array([0.5, 0.04, 0.05, 0.9, 0.9], np.float32)
    flags = np.array([GAS, FLUID, FLUID, SOLID, IFACE], np.uint8)
    fl = np.full(5, 0.7, np.float32)
    lib().or_sync_phi_liquid_fraction(ptr(phi), ptr(fl), ptr(flags), 5, 1, 1)
    assert fl.tolist() == [0.0, 0.0, np.float32(0.7), 0.0, np.float32(0.7)]


# ---------------------------------------------------------------------------- thermal
def test_boundary_losses_closed_form():
    """A FLUID cell on a 3x3x3 block with GAS above and on its +x side has two exposed faces
    (+z, +x; the -z face never counts): dH = -(h (Te - Ta) + eps sigma (Te^4 - Ta^4)) * 2 dt/dx,
    Te = min(T, 2500) (kernels.py:2030-2082)."""
    nx = ny = nz = 3
    flags = np.full((nx, ny, nz), FLUID, np.uint8)
    flags[:, :, 2] = GAS
    flags[2, :, :] = GAS
    T = np.full((nx, ny, nz), 300.0, np.float32)
    T[1, 1, 1] = 2900.0
    H = np.zeros((nx, ny, nz), np.float32)
    h, eps, sig, Ta, dt, dx = 25.0, 0.3, 5.670374419e-8, 300.0, 2e-5, 2e-4
    lib().or_apply_thermal_boundary_losses(ptr(H), ptr(T), ptr(flags), Ta, h, eps, 1, 1, 1.0, dt,
                                           dx, sig, nx, ny, nz)
    Te = 2500.0
    q = h * (Te - Ta) + eps * sig * (Te ** 4 - Ta ** 4)
    assert math.isclose(H[1, 1, 1], -q * 2 * dt / dx, rel_tol=2e-6)
    # an interior cell with no GAS neighbour loses nothing
    assert H[0, 0, 0] == 0.0 and H[1, 1, 0] == 0.0


def test_enthalpy_floor_and_ceiling():
    cp_rho, Ts, Tl, L, Tcap = 7800.0 * 680.0, 1723.0, 1780.0, 7800.0 * 2.7e5, 3200.0
    H_sol = cp_rho * Ts
    H_liq = H_sol + L
    cap = H_liq + cp_rho * (Tcap - Tl)
    H = np.array([0.0, cp_rho * 250.0, cap * 1.5, cap * 0.5, cap * 2.0], np.float32)
    flags = np.array([FLUID, SOLID, FLUID, FLUID, GAS], np.uint8)
    Hf = H.copy()
    lib().or_clamp_enthalpy_floor_scalar(ptr(Hf), cp_rho, ptr(flags), 300.0, 5, 1, 1)
    assert Hf[0] == np.float32(np.float32(cp_rho) * np.float32(300.0))
    assert Hf[1] == np.float32(np.float32(cp_rho) * np.float32(300.0))
    assert Hf[2] == H[2] and Hf[4] == H[4]
    Hc = H.copy()
    lib().or_clamp_enthalpy_ceiling_scalar(ptr(Hc), ptr(flags), cp_rho, Ts, Tl, Tcap, L, 5, 1, 1)
    assert math.isclose(Hc[2], cap, rel_tol=1e-6)       # capped at T = T_cap
    assert Hc[3] == H[3] and Hc[4] == H[4]              # below the cap / GAS untouched


# ---------------------------------------------------------------------------- forces
def test_arc_pressure_closed_form():
    """Fz -= P/(rho dx) dt^2/dx exp(-(di^2 + dj^2 + 0.1 dk^2)/(2 s^2 + 1e-6)) where
    |phi(k+1) - phi(k-1)| > 0.05 (kernels.py:3038-3083)."""
    nx, ny, nz = 5, 5, 5
    flags = np.full((nx, ny, nz), FLUID, np.uint8)
    phi = np.zeros((nx, ny, nz), np.float32)
    phi[:, :, :2] = 1.0                  # interface between k=1 and k=2 -> k = 1, 2 qualify
    Fz = np.zeros((nx, ny, nz), np.float32)
    ai, aj, ak, s, Pp, dt, dx, rho = 2.0, 1.5, 2.0, 1.7, 500.0, 2e-5, 2e-4, 7800.0
    lib().or_apply_arc_pressure(ptr(Fz), ptr(flags), ptr(phi), ai, aj, ak, s, Pp, dt, dx, rho,
                                nx, ny, nz)
    peak = Pp / (rho * dx) * dt * dt / dx
    for (i, j, k) in [(2, 2, 1), (0, 4, 2), (3, 1, 2)]:
        r2 = (i - ai) ** 2 + (j - aj) ** 2 + 0.1 * (k - ak) ** 2
        assert math.isclose(Fz[i, j, k], -peak * math.exp(-r2 / (2 * s * s + 1e-6)), rel_tol=3e-6)
    assert Fz[2, 2, 4] == 0.0 and Fz[2, 2, 0] == 0.0


def test_buoyancy_and_hydrostatic_closed_forms():
    T = np.array([1900.0, 1700.0, 2000.0], np.float32)
    fl = np.array([0.5, 0.5, 0.0], np.float32)
    flags = np.array([FLUID, IFACE, FLUID], np.uint8)
    Fz = np.zeros(3, np.float32)
    g, beta, Tl = 4.9e-5, 1.2e-4, 1780.0
    lib().or_add_buoyancy(ptr(T), ptr(Fz), ptr(fl), ptr(flags), g, beta, Tl, 1.0, 3, 1, 1)
    assert math.isclose(Fz[0], g * beta * (1900.0 - Tl) * 0.5, rel_tol=2e-6)
    assert math.isclose(Fz[1], g * beta * (1700.0 - Tl) * 0.5, rel_tol=2e-6)
    assert Fz[2] == 0.0
    Fh = np.zeros(3, np.float32)
    lib().or_add_hydrostatic_gravity(ptr(Fh), ptr(fl), ptr(flags), 1.0, g, 3, 1, 1)
    assert math.isclose(Fh[0], -g * 0.5, rel_tol=1e-6) and Fh[2] == 0.0


def test_force_and_velocity_clamps_keep_direction():
    rng = np.random.default_rng(8)
    v = rng.normal(size=(3, 200)).astype(np.float32) * 0.1
    flags = np.full(200, FLUID, np.uint8)
    flags[:5] = SOLID
    x, y, z = (np.ascontiguousarray(v[a]) for a in range(3))
    lib().or_clamp_body_force_magnitude(ptr(x), ptr(y), ptr(z), ptr(flags), 0.05, 200, 1, 1)
    m0 = np.sqrt((v.astype(np.float64) ** 2).sum(0))
    m1 = np.sqrt(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2 + z.astype(np.float64) ** 2)
    big = (m0 > 0.05) & (flags != SOLID)
    assert np.all(m1[big] <= 0.05 * (1 + 1e-6)) and np.allclose(m1[big], 0.05, rtol=1e-6)
    cosang = (x * v[0] + y * v[1] + z * v[2]) / (m0 * m1)
    assert np.allclose(cosang[big], 1.0, atol=1e-6)
    assert np.array_equal(x[~big], v[0][~big])


# ---------------------------------------------------------------------------- VOF rules
def _contact(phi_col, flags_col, theta_deg, lateral=None):
    """Jacobi contact-angle BC on a 3x3 stack of copies of one column (lateral: phi of the
    side columns at the centre's level)."""
    nz = len(phi_col)
    phi = np.tile(np.asarray(phi_col, np.float32), (3, 3, 1))
    flags = np.tile(np.asarray(flags_col, np.uint8), (3, 3, 1))
    if lateral is not None:
        k, val = lateral
        phi[:, :, k] = 0.0
        phi[0, 1, k] = val
    out = phi.copy()
    t = math.radians(theta_deg)
    lib().or_apply_contact_angle_phi_bc_jacobi(ptr(phi), ptr(out), ptr(flags),
                                               np.float32(math.cos(t)), np.float32(math.sin(t)),
                                               3, 3, nz)
    return out[1, 1]


def test_contact_angle_rules():
    """Rules of kernels.py:949-991 with film = (1 - cos t)/2, on hand-built columns."""
    t = 120.0
    film = 0.5 * (1 - math.cos(math.radians(t)))   # 0.75
    s = math.sin(math.radians(t))
    # rule 1: a GAS cell right above SOLID holds at least the film
    out = _contact([1.0, 0.0, 0.0], [SOLID, GAS, GAS], t)
    assert math.isclose(out[1], film, rel_tol=1e-6)
    # rule 3: GAS above a mostly-liquid cell: max(p, clamp(phi_below - sin t, film, 0.95))
    out = _contact([1.0, 1.0, 0.9, 0.0], [SOLID, FLUID, FLUID, GAS], t)
    assert math.isclose(out[3], max(film, min(0.9 - s, 0.95)), rel_tol=1e-6)
    # rule 5: a wetted cell on SOLID with phi in (0.2, 0.65) is pinned to 0.65
    out = _contact([1.0, 0.3, 0.0], [SOLID, IFACE, GAS], 30.0)
    assert math.isclose(out[1], 0.65, rel_tol=1e-6)
    # rule 4: lateral spreading into an empty cell next to a liquid one
    film30 = 0.5 * (1 - math.cos(math.radians(30.0)))
    out = _contact([1.0, 1.0, 0.0, 0.0], [SOLID, FLUID, IFACE, GAS], 30.0, lateral=(2, 0.9))
    assert math.isclose(out[2], film30 * (0.85 + 0.15 * math.sin(math.radians(30.0))),
                        rel_tol=1e-6)
    # SOLID cells are never changed
    out = _contact([0.3, 0.0], [SOLID, GAS], t)
    assert out[0] == np.float32(0.3)


def test_film_threshold_identity():
    """film > 0.2 exactly when theta > 53.13 deg (cos theta < 0.6)."""
    for deg, above in [(53.0, False), (53.2, True), (90.0, True), (30.0, False)]:
        film = 0.5 * (1 - math.cos(math.radians(deg)))
        assert (film > 0.2) == above


# ---------------------------------------------------------------------------- freeze rules
def test_cooled_metal_and_remelt_rules():
    """solidify_cooled_metal: phi >= 0.85, f_l < 0.02, T < T_s -> SOLID, phi 1, f_l 0, u 0;
    remelt_hot_solid_scalar: SOLID with T > T_s + 5 -> FLUID, phi 1, lever-rule f_l and H
    (kernels.py:2422-2450, 2489-2520)."""
    Ts, Tl, cp_rho, L = 1723.0, 1780.0, 5.304e6, 2.106e9
    T = np.array([1700.0, 1700.0, 1725.0, 1750.0, 1900.0], np.float32)
    fl = np.array([0.01, 0.03, 0.01, 0.0, 0.0], np.float32)
    phi = np.array([0.9, 0.9, 0.9, 1.0, 1.0], np.float32)
    flags = np.array([FLUID, FLUID, FLUID, SOLID, SOLID], np.uint8)
    ux = np.full(5, 0.01, np.float32)
    uy, uz = ux.copy(), ux.copy()
    f2, p2, fg2 = fl.copy(), phi.copy(), flags.copy()
    lib().or_solidify_cooled_metal(ptr(T), ptr(f2), ptr(p2), ptr(fg2), ptr(ux), ptr(uy), ptr(uz),
                                   Ts, 1, 5, 1, 1)
    assert fg2[0] == SOLID and p2[0] == 1.0 and f2[0] == 0.0 and ux[0] == 0.0
    assert fg2[1] == FLUID and fg2[2] == FLUID and ux[1] == np.float32(0.01)
    H = np.zeros(5, np.float32)
    lib().or_remelt_hot_solid_scalar(ptr(T), ptr(H), ptr(fl), ptr(phi), ptr(flags), cp_rho, Ts,
                                     Tl, L, 5, 1, 1)
    assert flags[3] == FLUID and flags[4] == FLUID and phi[3] == 1.0
    assert math.isclose(fl[3], (1750.0 - Ts) / (Tl - Ts + 1e-6), rel_tol=1e-6)
    assert math.isclose(H[3], cp_rho * Ts + fl[3] * L, rel_tol=1e-6)
    assert fl[4] == 1.0 and math.isclose(H[4], cp_rho * Tl + L + cp_rho * (1900.0 - Tl),
                                         rel_tol=1e-6)
    assert flags[0] == FLUID and H[0] == 0.0   # not SOLID: remelt leaves it alone


# ---------------------------------------------------------------------------- telemetry
def test_telemetry_matches_brute_force():
    """Pool aggregates of kernels.py:3424-3468 recomputed with numpy on a random state."""
    nx, ny, nz = 7, 6, 5
    rng = np.random.default_rng(12)
    T = rng.uniform(300, 3000, (nx, ny, nz)).astype(np.float32)
    fl = rng.uniform(0, 1, (nx, ny, nz)).astype(np.float32)
    phi = rng.uniform(0, 1, (nx, ny, nz)).astype(np.float32)
    dT = rng.uniform(-5e5, 5e5, (nx, ny, nz)).astype(np.float32)
    u = [rng.uniform(-0.05, 0.05, (nx, ny, nz)).astype(np.float32) for _ in range(3)]
    out = np.zeros(8, np.float64)
    lib().or_telemetry_pool_reduce(ptr(T), ptr(fl), ptr(phi), ptr(dT), ptr(u[0]), ptr(u[1]),
                                   ptr(u[2]), ptr(out), nx, ny, nz)
    liq = fl > 0.5
    ii = np.broadcast_to(np.arange(nx)[:, None, None], T.shape)
    u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2]
    assert out[0] == liq.sum()
    assert out[1] == T[liq].max()
    assert out[2] == max(0.0, float((-dT[phi > 0.05]).max()))
    assert out[3] == u2[liq].max()
    assert out[4] == ii[liq].sum() and out[5] == ii[liq].min() and out[6] == ii[liq].max()
    assert out[7] == T.max()


# ---------------------------------------------------------------------------- deposition
def _plate(nx=16, ny=16, nz=12, k0=4):
    flags = np.full((nx, ny, nz), GAS, np.uint8)
    flags[:, :, :k0] = SOLID
    phi = np.where(flags == GAS, 0.0, 1.0).astype(np.float32)
    f_l = np.zeros(flags.shape, np.float32)
    T = np.full(flags.shape, 300.0, np.float32)
    H = (T * np.float32(5.3e6)).astype(np.float32)
    rho = np.ones(flags.shape, np.float32)
    f = np.zeros((19,) + flags.shape, np.float32)
    return dict(flags=flags, phi=phi, f_l=f_l, T=T, H=H, rho=rho, f=f)


def _feed(s, arc, foot, drop_r, target, T_drop=2200.0, cp=5.3e6, L=2.1e9):
    acc = np.zeros(1, np.float32)
    real = np.zeros(1, np.float32)
    lib().or_feed_wire_surface(ptr(s["f"]), ptr(s["flags"]), ptr(s["f_l"]), ptr(s["phi"]),
                               ptr(s["H"]), ptr(s["T"]), ptr(s["rho"]), *arc, foot, drop_r,
                               target, T_drop, cp, L, 1.0, ptr(acc), ptr(real), 1.0,
                               *s["flags"].shape)
    return float(acc[0]), float(real[0])


def test_deposition_serial_sweep_builds_columns_within_budget():
    """A serial i, j, k sweep with a metal-neighbour test deposits whole columns: on a flat
    plate (top at k0 = 4) each footprint column (i-8)^2 + (j-8)^2 <= 9, in lexicographic (i, j)
    order, fills k = 4 .. k_hi - 1 with k_hi = int(arc_k + max(r_d + 3, 4)) + 1 = 10, because
    every new cell is the metal neighbour of the one above it. Reserve-then-commit against the
    budget stops after ceil(7.5 / 1) = 8 cells: column (5, 8) [6 cells], then (6, 6, 4..5).
    Pass 3 re-heats existing melt (f_l >= 0.45) within 1.2 x the footprint only."""
    s = _plate()
    s["flags"][8, 8, 3] = FLUID
    s["f_l"][8, 8, 3] = 0.5
    s["flags"][1, 1, 3] = FLUID
    s["f_l"][1, 1, 3] = 0.5
    before = s["flags"].copy()
    acc, real = _feed(s, (8.0, 8.0, 4.0), 3.0, 2.0, 7.5)
    new = np.argwhere((before == GAS) & (s["flags"] == FLUID))
    expect = [(5, 8, k) for k in range(4, 10)] + [(6, 6, 4), (6, 6, 5)]
    assert sorted(map(tuple, new)) == sorted(expect)
    assert acc == 8.0 and real == 8.0
    cp, L = np.float32(5.3e6), np.float32(2.1e9)
    for c in expect:
        assert s["phi"][c] == 1.0 and s["f_l"][c] == 1.0 and s["T"][c] == np.float32(2200.0)
        assert s["H"][c] == cp * np.float32(2200.0) + L
        np.testing.assert_array_equal(s["f"][(slice(None),) + c], np.asarray(W, np.float32))
    assert s["T"][8, 8, 3] == np.float32(2200.0) and s["T"][1, 1, 3] == np.float32(300.0)


def test_deposition_stops_when_budget_is_met_and_never_without_metal():
    s = _plate()
    acc, real = _feed(s, (8.0, 8.0, 4.0), 3.0, 2.0, 1000.0)
    new = (s["flags"] == FLUID)
    # unlimited budget: every footprint column up to k_hi - 1, nothing outside the footprint
    ii, jj = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    col = ((ii - 8) ** 2 + (jj - 8) ** 2) <= 9 * 1.44
    assert np.all(new[~col] == False)  # noqa: E712
    assert real == float(new.sum()) == acc
    s2 = _plate()
    s2["flags"][:] = GAS
    s2["phi"][:] = 0.0
    acc, real = _feed(s2, (8.0, 8.0, 4.0), 3.0, 2.0, 10.0)
    assert acc == 0.0 and real == 0.0 and np.all(s2["flags"] == GAS)


def test_momentum_impact_sets_equilibrium_at_the_droplet_velocity():
    """Cells in the droplet region get u = v and f = f_eq(rho, v): sum f = rho and
    sum f e = rho v (D3Q19 moment identities); GAS and mushy (f_l < 0.45) cells are untouched."""
    rng = np.random.default_rng(9)
    n = (12, 12, 12)
    flags = np.full(n, FLUID, np.uint8)
    f_l = np.ones(n, np.float32)
    flags[6, 6, 7] = GAS
    f_l[5, 6, 6] = 0.3
    rho = rng.uniform(0.95, 1.05, n).astype(np.float32)
    u = [np.zeros(n, np.float32) for _ in range(3)]
    f = rng.uniform(0.0, 0.1, (19,) + n).astype(np.float32)
    f0 = f.copy()
    v = (0.01, -0.02, -0.05)
    lib().or_feed_wire_momentum_impact(ptr(f), ptr(rho), *(ptr(a) for a in u), ptr(flags),
                                       ptr(f_l), 6.0, 6.0, 3.0, 1.5, *v, *n)
    hit = u[2] != 0
    assert hit[6, 6, 5] and not hit[6, 6, 7] and not hit[5, 6, 6] and not hit[0, 0, 0]
    for a, va in zip(u, v):
        assert np.all(a[hit] == np.float32(va))
    fs = f[:, hit].astype(np.float64)
    np.testing.assert_allclose(fs.sum(0), rho[hit], rtol=2e-6)
    for a, va in zip(E, v):
        np.testing.assert_allclose((a[:, None] * fs).sum(0), rho[hit] * va, rtol=2e-5, atol=1e-8)
    np.testing.assert_array_equal(f[:, ~hit], f0[:, ~hit])


def test_droplet_impact_pressure_closed_form():
    """Fz += -p/(rho dx) dt^2/dx exp(-r^2 / (2 r_d^2)) on metal or wetted cells within 2 r_d."""
    n = (14, 14, 10)
    flags = np.full(n, FLUID, np.uint8)
    flags[:, :, 8:] = GAS
    phi = np.ones(n, np.float32)
    f_l = np.ones(n, np.float32)
    f_l[3, 7, 4], phi[3, 7, 4] = 0.1, 0.1
    Fz = np.zeros(n, np.float32)
    p, dt, dx, rho_ref, dr = 800.0, 2e-5, 2e-4, 7800.0, 2.0
    lib().or_apply_droplet_impact_pressure(ptr(Fz), ptr(flags), ptr(phi), ptr(f_l), 7.0, 7.0, 4.0,
                                           dr, p, dt, dx, rho_ref, *n)
    i, j, k = np.meshgrid(*(np.arange(m) for m in n), indexing="ij")
    r2 = (i - 7.0) ** 2 + (j - 7.0) ** 2 + (k - 4.0) ** 2
    expect = -p / (rho_ref * dx) * dt * dt / dx * np.exp(-r2 / (2 * dr * dr))
    on = (r2 <= 4 * dr * dr) & (flags != GAS)
    on[3, 7, 4] = False
    np.testing.assert_allclose(Fz[on], expect[on], rtol=2e-6)
    assert np.all(Fz[~on] == 0.0)


# ---------------------------------------------------------------------------- tracers
def test_tracer_ring_buffer_and_spawn_box():
    """Slots wrap modulo max_tracers (the oldest are evicted); spawns sit in the +/- sigma box
    around (x, y) at height z."""
    pos = np.zeros((5, 3), np.float32)
    active = np.zeros(5, np.int32)
    head = np.zeros(1, np.int32)
    lib().or_inject_tracers(ptr(pos), ptr(active), ptr(head), 5, 1e-3, 2e-3, 3e-4, 2e-4, 7, 11)
    assert head[0] == 7 and np.all(active == 1)
    assert np.all(np.abs(pos[:, 0] - np.float32(1e-3)) <= 2e-4 * (1 + 1e-6))
    assert np.all(np.abs(pos[:, 1] - np.float32(2e-3)) <= 2e-4 * (1 + 1e-6))
    assert np.all(pos[:, 2] == np.float32(3e-4))
    assert len(set(map(tuple, pos[:, :2].tolist()))) == 5


def test_tracer_advection_states():
    """A tracer in melt moves u dx per step (u in cells/step); entering GAS or leaving the domain
    deactivates it (0); SOLID or f_l < 0.05 traps it for good (2)."""
    n = (6, 5, 4)
    dx = 1e-4
    flags = np.full(n, FLUID, np.uint8)
    flags[0, 0, 0] = GAS
    flags[1, 0, 0] = SOLID
    f_l = np.ones(n, np.float32)
    f_l[2, 0, 0] = 0.01
    ux = np.full(n, 0.25, np.float32)
    uy = np.full(n, -0.125, np.float32)
    uz = np.full(n, 0.5, np.float32)
    cells = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 2, 1), (9, 0, 0)]
    pos = np.array([[(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx] for i, j, k in cells],
                   np.float32)
    active = np.array([1, 1, 1, 1, 1], np.int32)
    p0 = pos.copy()
    lib().or_advect_tracers(ptr(pos), ptr(active), ptr(ux), ptr(uy), ptr(uz), ptr(f_l),
                            ptr(flags), dx, 2e-5, 5, *n)
    assert list(active) == [0, 2, 2, 1, 0]
    np.testing.assert_array_equal(pos[[0, 1, 2, 4]], p0[[0, 1, 2, 4]])
    moved = p0[3] + np.array([0.25, -0.125, 0.5], np.float32) * np.float32(dx)
    np.testing.assert_array_equal(pos[3], moved)


def test_trailing_pool_rules_at_their_thresholds():
    """solidify_trailing_pool_scalar (kernels.py:2574-2617): a non-gas cell with phi >= 0.85,
    f_l >= 0.05, projection on the torch direction <= -lookback and T < T_freeze freezes to
    T_s - 1 (H = cp_rho (T_s - 1)), SOLID, phi 1, f_l 0, u 0; each threshold is inclusive on the
    freezing side and exclusive on the other."""
    Ts, Tf, cp = 1723.0, 1800.0, 5.3e6
    # cells along x at k = 0; arc at i = 10 moving +x, lookback 3 -> i <= 7 is behind
    rows = [  # (i, flag, phi, f_l, T, freezes)
        (7, FLUID, 0.85, 0.05, 1799.0, True),     # every threshold at its boundary
        (8, FLUID, 1.0, 1.0, 1700.0, False),      # proj = -2 > -3
        (6, FLUID, 0.84, 1.0, 1700.0, False),     # phi below 0.85
        (5, FLUID, 1.0, 0.049, 1700.0, False),    # f_l below 0.05
        (4, FLUID, 1.0, 1.0, 1800.0, False),      # T == T_freeze
        (3, GAS, 1.0, 1.0, 1700.0, False),        # gas never freezes
        (2, IFACE, 0.9, 0.5, 1000.0, True),
        (1, SOLID, 1.0, 0.5, 1000.0, True),       # any non-gas flag
    ]
    n = (12, 1, 1)
    flags = np.full(n, GAS, np.uint8)
    phi, f_l, T = (np.zeros(n, np.float32) for _ in range(3))
    for i, fg, p, fl, t, _ in rows:
        flags[i, 0, 0], phi[i, 0, 0], f_l[i, 0, 0], T[i, 0, 0] = fg, p, fl, t
    H = np.full(n, 7.0, np.float32)
    u = [np.full(n, 0.01, np.float32) for _ in range(3)]
    lib().or_solidify_trailing_pool_scalar(ptr(T), ptr(H), ptr(f_l), ptr(phi), ptr(flags),
                                           *(ptr(a) for a in u), cp, 10.0, 0.0, 0.0, 1.0, 0.0,
                                           0.0, 3.0, Tf, Ts, *n)
    for i, fg, p, fl, t, freezes in rows:
        c = (i, 0, 0)
        if freezes:
            assert flags[c] == SOLID and phi[c] == 1.0 and f_l[c] == 0.0
            assert T[c] == np.float32(Ts - 1.0)
            assert H[c] == np.float32(cp) * (np.float32(Ts) - np.float32(1.0))
            assert u[0][c] == 0.0 and u[2][c] == 0.0
        else:
            assert flags[c] == fg and T[c] == np.float32(t) and H[c] == 7.0


# ---------------------------------------------------------------------------- moving window
@pytest.mark.parametrize("n_shift", [3, 8, 12, 20])
def test_window_shift_is_a_pure_shift_with_reset_strip(n_shift):
    """An ascending serial loop that copies cell i + n into cell i reads only cells not yet
    written, so every field becomes new[:nx-n] = old[n:] (a descending loop would re-read
    shifted cells); the last n columns take the ambient plate/gas state (kernels.py:2620-2723).
    n >= nx resets the whole domain."""
    from oracle.oracle import OracleTwin

    nx, ny, nz = 12, 3, 5
    rng = np.random.default_rng(n_shift)
    names = OracleTwin.WINDOW_FIELDS
    fl = [rng.standard_normal((nx, ny, nz)).astype(np.float32) for _ in names]
    flags = rng.integers(0, 5, (nx, ny, nz)).astype(np.uint8)
    f_a = rng.standard_normal((19, nx, ny, nz)).astype(np.float32)
    f_b = rng.standard_normal((19, nx, ny, nz)).astype(np.float32)
    old = [a.copy() for a in fl], flags.copy(), f_a.copy(), f_b.copy()
    import ctypes

    table = (ctypes.c_void_p * len(fl))(*[a.ctypes.data for a in fl])
    lib().or_shift_simulation_window_x(n_shift, ptr(f_a), ptr(f_b), table, ptr(flags), 293.0,
                                       5.3e6, 1.0, 2, nx, ny, nz)
    m = max(0, nx - n_shift)
    for a, o in zip(fl + [flags], old[0] + [old[1]]):
        np.testing.assert_array_equal(a[:m], o[n_shift:n_shift + m])
    for a, o in ((f_a, old[2]), (f_b, old[3])):
        np.testing.assert_array_equal(a[:, :m], o[:, n_shift:n_shift + m])
        np.testing.assert_array_equal(a[:, m:], np.broadcast_to(
            np.asarray(W, np.float32)[:, None, None, None], a[:, m:].shape))
    d = dict(zip(names, fl))
    strip = slice(m, nx)
    assert np.all(d["T"][strip] == np.float32(293.0))
    assert np.all(d["H"][strip] == np.float32(5.3e6) * np.float32(293.0))
    assert np.all(d["T_max"][strip] == np.float32(293.0)) and np.all(d["rho"][strip] == 1.0)
    for n in ("f_l", "ux", "uy", "uz", "Fx", "Fy", "Fz", "dT_dt", "time_above_800_s",
              "time_above_1100_s", "time_above_solidus_s"):
        assert np.all(d[n][strip] == 0.0), n
    # the per-cell material fields are shifted but not reset
    for n in ("cp_rho_field", "alpha_lu_field", "dgamma_lu_field", "tau_field"):
        np.testing.assert_array_equal(d[n][strip], old[0][names.index(n)][strip])
    assert np.all(flags[strip][:, :, :2] == SOLID) and np.all(flags[strip][:, :, 2:] == GAS)
    assert np.all(d["phi"][strip][:, :, :2] == 1.0) and np.all(d["phi"][strip][:, :, 2:] == 0.0)


# ---------------------------------------------------------------------------- full-tier forces
def _surface_column(T_val, n=(9, 7, 6), kz=3):
    """A FLUID layer at k <= kz under GAS, so |phi(k+1) - phi(k-1)| = 1 at the top metal cell."""
    flags = np.full(n, GAS, np.uint8)
    flags[:, :, :kz + 1] = FLUID
    phi = np.where(flags == GAS, 0.0, 1.0).astype(np.float32)
    T = np.full(n, T_val, np.float32)
    return flags, phi, T


@pytest.mark.parametrize("T_val", [1500.0, 2600.0, 9000.0])
def test_constant_recoil_closed_form(T_val):
    """Fz = -(p / (rho dx)) dt^2 / dx min(T / T_ref, 3) exp(-r^2 / (2 sigma^2)) at the surface
    cells (|d phi / dz| >= 0.05), nothing elsewhere (kernels.py:995-1030)."""
    flags, phi, T = _surface_column(T_val)
    Fz = np.zeros(T.shape, np.float32)
    p, T_ref, dt, dx, rho, sig, ai, aj = 5000.0, 1780.0, 2e-5, 2e-4, 7800.0, 3.0, 4.0, 3.0
    lib().or_apply_vapor_recoil(ptr(Fz), ptr(T), ptr(phi), ptr(flags), ai, aj, sig, p, T_ref, dt,
                                dx, rho, *T.shape)
    i, j = np.meshgrid(np.arange(T.shape[0]), np.arange(T.shape[1]), indexing="ij")
    r2 = (i - ai) ** 2 + (j - aj) ** 2
    expect = -(p / (rho * dx)) * dt * dt / dx * min(T_val / T_ref, 3.0) * np.exp(-r2 / (2 * sig ** 2))
    np.testing.assert_allclose(Fz[:, :, 3], expect, rtol=2e-6)
    assert np.all(Fz[:, :, :3] == 0) and np.all(Fz[:, :, 4:] == 0)


def test_clausius_clapeyron_recoil_closed_form():
    """p = C_acc P_ref exp(min(L/R (1/T_b - 1/T), 12)) above the boiling point, zero at or below
    it (kernels.py:1052-1090)."""
    for T_val in (3000.0, 3100.0, 3400.0, 20000.0):
        flags, phi, T = _surface_column(T_val)
        Fz = np.zeros(T.shape, np.float32)
        P_ref, Tb, L, R, C, dt, dx, rho, sig = 101325.0, 3100.0, 6e6, 450.0, 0.54, 2e-5, 2e-4, 7800.0, 3.0
        lib().or_apply_vapor_recoil_cc(ptr(Fz), ptr(T), ptr(phi), ptr(flags), 4.0, 3.0, sig, P_ref,
                                       Tb, L, R, C, dt, dx, rho, *T.shape)
        if T_val <= Tb:
            assert np.all(Fz == 0)
            continue
        P = C * P_ref * math.exp(min(L / R * (1 / Tb - 1 / T_val), 12.0))
        F0 = P / (rho * dx) * dt * dt / dx
        assert Fz[4, 3, 3] == pytest.approx(-F0, rel=2e-5)
        assert Fz[4, 5, 3] == pytest.approx(-F0 * math.exp(-4 / 18), rel=2e-5)


def test_gas_shear_is_tangential_and_radial():
    """On a flat surface the shear is horizontal, points away from the arc axis and has magnitude
    tau exp(-r^2 / (2 sigma^2)) / (rho dx) dt^2 / dx (kernels.py:1093-1139); cells with
    |grad phi| < 0.08 get none."""
    flags, phi, T = _surface_column(1800.0)
    Fx, Fy = np.zeros(T.shape, np.float32), np.zeros(T.shape, np.float32)
    tau, dt, dx, rho, sig, ai, aj = 115.2, 2e-5, 2e-4, 7800.0, 3.0, 4.0, 3.0
    lib().or_apply_gas_shear_stress(ptr(Fx), ptr(Fy), ptr(phi), ptr(flags), ai, aj, sig, tau, dt,
                                    dx, rho, *T.shape)
    scale = 1.0 / (rho * dx) * dt * dt / dx
    for i in range(T.shape[0]):
        for j in range(T.shape[1]):
            di, dj = i - ai, j - aj
            r2 = di * di + dj * dj
            mag = math.hypot(Fx[i, j, 3], Fy[i, j, 3])
            assert mag == pytest.approx(tau * math.exp(-r2 / (2 * sig ** 2)) * scale *
                                        math.sqrt(r2 / (r2 + 1e-6)) if r2 > 0 else mag,
                                        rel=1e-4, abs=1e-30)
            if r2 > 0:
                assert Fx[i, j, 3] * di + Fy[i, j, 3] * dj > 0
    # interior cells (grad phi = 0) and gas get nothing
    assert np.all(Fx[:, :, :2] == 0) and np.all(Fx[:, :, 5:] == 0)


# ---------------------------------------------------------------------------- Lorentz force
def _elec_grid(n=(14, 12, 10), nzs=3):
    flags = np.full(n, FLUID, np.uint8)
    flags[:, :, :nzs] = SOLID
    flags[:, :, n[2] - 2:] = GAS
    f_l = np.where(flags == FLUID, 1.0, 0.0).astype(np.float32)
    return flags, f_l


def test_current_source_carries_the_welding_current():
    """After normalisation, sum(source) dx^3 equals the welding current (closed form)."""
    flags, f_l = _elec_grid()
    src = np.zeros(flags.shape, np.float32)
    dx = 2e-4
    lib().or_elec_inject_arc_source(ptr(src), 7.0, 6.0, 7.0, 2.0, 140.0, dx, *flags.shape)
    assert np.all(src[:, :, :2] == 0)               # below arc_k - sigma: untouched
    lib().or_elec_normalize_source(ptr(src), 140.0, dx, *flags.shape)
    assert float(np.sum(src.astype(np.float64)) * dx ** 3) == pytest.approx(140.0, rel=1e-5)


def test_jacobi_converges_to_the_discrete_conduction_equation():
    """The fixed point of the weighted Jacobi sweep solves sum_f sigma_f (phi_f - phi) +
    s dx^2 = 0 on every non-grounded metal cell (the defining discrete equation), with phi = 0 on
    GAS and on SOLID at k <= nz_solid."""
    flags, f_l = _elec_grid()
    n = flags.shape
    L = lib()
    sigma = np.zeros(n, np.float32)
    L.or_elec_build_sigma(ptr(sigma), ptr(f_l), ptr(flags), 1e6, 1e5, *n)
    assert sigma[0, 0, 0] == np.float32(1e5) and sigma[5, 5, 5] == np.float32(1e6)
    assert sigma[0, 0, n[2] - 1] == np.float32(1e-8)
    src = np.zeros(n, np.float32)
    dx = 2e-4
    L.or_elec_inject_arc_source(ptr(src), 7.0, 6.0, 7.0, 2.0, 140.0, dx, *n)
    L.or_elec_normalize_source(ptr(src), 140.0, dx, *n)
    a, b = np.zeros(n, np.float32), np.zeros(n, np.float32)
    for _ in range(6000):
        L.or_elec_jacobi_step(ptr(a), ptr(b), ptr(sigma), ptr(src), ptr(flags), dx, 0.85, 3, *n)
        a, b = b, a
    s = sigma.astype(np.float64)
    phi = a.astype(np.float64)
    pad = lambda x: np.pad(x, 1, mode="edge")  # noqa: E731  clamped neighbours
    sp, pp = pad(s), pad(phi)
    res = np.zeros(n)
    denom = np.zeros(n)
    for sl in ((slice(2, None), slice(1, -1), slice(1, -1)), (slice(0, -2), slice(1, -1), slice(1, -1)),
               (slice(1, -1), slice(2, None), slice(1, -1)), (slice(1, -1), slice(0, -2), slice(1, -1)),
               (slice(1, -1), slice(1, -1), slice(2, None)), (slice(1, -1), slice(1, -1), slice(0, -2))):
        res += sp[sl] * (pp[sl] - phi)
        denom += sp[sl]
    res += src * dx * dx
    k = np.arange(n[2])[None, None, :]
    active = (flags != GAS) & ~((k <= 3) & (flags == SOLID))
    assert np.all(phi[~active] == 0)
    scale = np.max(np.abs(src * dx * dx))
    assert np.max(np.abs(res[active] / denom[active])) < 1e-4 * scale / np.max(denom)
    out = np.zeros(2, np.float32)
    L.or_elec_jacobi_step(ptr(a), ptr(b), ptr(sigma), ptr(src), ptr(flags), dx, 0.85, 3, *n)
    L.or_elec_l1_diff(ptr(b), ptr(a), ptr(out), *n)
    assert out[0] / out[1] < 1e-6 and out[1] == pytest.approx(np.abs(a).astype(np.float64).sum(), rel=1e-6)


def test_current_density_and_self_field_closed_forms():
    """J = -sigma grad(phi) for a linear potential; for a uniform axial current density the
    enclosed current of each radial bin is Jz dx^2 times the cells inside it, and
    B_theta = mu0 I_enc / (2 pi r) circles the axis; J x B points inward (pinch)."""
    n = (16, 16, 4)
    flags = np.full(n, FLUID, np.uint8)
    f_l = np.ones(n, np.float32)
    sigma = np.full(n, 2.0e5, np.float32)
    dx = 2e-4
    k = np.arange(n[2], dtype=np.float32)[None, None, :]
    phi = np.broadcast_to(-0.01 * k, n).astype(np.float32).copy()
    J = [np.zeros(n, np.float32) for _ in range(3)]
    L = lib()
    L.or_elec_compute_J(*(ptr(x) for x in J), ptr(phi), ptr(sigma), ptr(flags), dx, *n)
    # interior k: central difference of -0.01 k is -0.01 per cell
    np.testing.assert_allclose(J[2][:, :, 1:3], 2.0e5 * 0.01 / dx, rtol=1e-6)
    assert np.all(J[0] == 0) and np.all(J[1] == 0)
    bins = np.zeros((n[2], 64), np.float32)
    dr = 0.5
    L.or_elec_bin_axial_current(ptr(J[2]), ptr(bins), 7.5, 7.5, dx, dr, 64, *n)
    ii, jj = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    r = np.sqrt((ii - 7.5) ** 2 + (jj - 7.5) ** 2)
    b = np.minimum((r / dr).astype(int), 63)
    cnt = np.bincount(b.ravel(), minlength=64)
    np.testing.assert_allclose(bins[1], J[2][0, 0, 1] * dx * dx * cnt, rtol=1e-6)
    L.or_elec_prefix_bins(ptr(bins), 64, n[2])
    np.testing.assert_allclose(bins[1], np.cumsum(J[2][0, 0, 1] * dx * dx * cnt), rtol=1e-5)
    B = [np.zeros(n, np.float32) for _ in range(3)]
    mu0 = 4e-7 * math.pi
    L.or_elec_B_axisymmetric(*(ptr(x) for x in B), ptr(bins), 7.5, 7.5, dx, dr, mu0, 64, *n)
    i, j = 12, 7
    rr = math.hypot(i - 7.5, j - 7.5)
    Bt = mu0 * bins[1, min(int(rr / dr), 63)] / (2 * math.pi * rr * dx)
    assert B[0][i, j, 1] == pytest.approx(-Bt * (j - 7.5) / rr, rel=1e-5)
    assert B[1][i, j, 1] == pytest.approx(Bt * (i - 7.5) / rr, rel=1e-5)
    F = [np.zeros(n, np.float32) for _ in range(3)]
    L.or_apply_lorentz_JxB(*(ptr(x) for x in F), *(ptr(x) for x in J), *(ptr(x) for x in B),
                           ptr(f_l), ptr(flags), 2e-5, dx, 7800.0, *n)
    # current along +z with B circling +z: the force points towards the axis
    assert F[0][i, j, 1] * (i - 7.5) + F[1][i, j, 1] * (j - 7.5) < 0
