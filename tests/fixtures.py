"""Host-side fixtures shared by the CPU and GPU tests.

Each builds float32 numpy arrays exactly as the corresponding reference fixture does, so the
same bytes can be handed to the oracle and uploaded to the GPU:
  * material constants of the reference's placeholder ER70S-6 file
    (reference materials/placeholders/ER70S-6.yaml) and the low-carbon-steel substrate of the
    256x128x64 bead-on-plate workload;
  * Poiseuille channel      — reference kernels.py:2086-2136 (init_poiseuille_channel);
  * Gaussian thermal pulse  — reference kernels.py:3089-3116 (prescribe_gaussian_pulse);
  * 1-D Stefan column       — reference kernels.py:3125-3157 (init_stefan_liquid_column);
  * tanh droplet            — reference validation/test_laplace.py:20-28;
  * seeded molten patch     — reference tools/force_ablation.py:71-111.
"""

from __future__ import annotations

import math

import numpy as np

FLUID, SOLID, GAS, IFACE = 0, 1, 2, 4
EX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
EY = (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1)
EZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1)
W32 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)

# Constants block of the reference's placeholder ER70S-6 material file.
ER70S6_PLACEHOLDER = dict(name="ER70S-6", k=30.0, rho=7000.0, cp=680.0, alpha=6.3e-6,
                          T_solidus=1748.0, T_liquidus=1793.0, L_fusion=272000.0, gamma_0=1.8,
                          dgamma_dT=-4.3e-4, mu=0.0065, beta_T=1.2e-4)

# Low-carbon-steel substrate of the 256x128x64 bead-on-plate workload (rho, cp, k, T_s, T_l,
# L and dgamma/dT as specified for it; gamma_0, mu and beta as in the ER70S-6 file).
LCS_BEAD_ON_PLATE = dict(name="LCS-bead-on-plate", k=30.0, rho=7800.0, cp=680.0,
                         T_solidus=1723.0, T_liquidus=1780.0, L_fusion=2.7e5, gamma_0=1.8,
                         dgamma_dT=-4.0e-4, mu=0.0065, beta_T=1.2e-4)

# Piecewise-linear knots of the reference's validated ER70S-6 v1 file (cp, k, mu) with its
# two-knot dgamma/dT table (the file's own surfactant model replaces that table; these raw knots
# keep the table lookups at three and two knots).
ER70S6_V1_TABLES = dict(cp=[(300.0, 650.0), (1000.0, 680.0), (1793.0, 720.0)],
                        k=[(300.0, 28.0), (1000.0, 30.0), (1793.0, 34.0)],
                        mu=[(300.0, 0.0075), (1793.0, 0.0065), (2000.0, 0.0055)],
                        dgamma_dT=[(1793.0, -4.3e-4), (2000.0, -3.8e-4)])


def f32(x) -> np.float32:
    return np.float32(x)


def equilibrium(rho: np.ndarray, ux: np.ndarray, uy: np.ndarray, uz: np.ndarray) -> np.ndarray:
    """W_q*rho*(1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2) in fp32, reference operand order."""
    u2 = ux * ux + uy * uy + uz * uz
    f = np.empty((19,) + rho.shape, np.float32)
    for q in range(19):
        eu = f32(EX[q]) * ux + f32(EY[q]) * uy + f32(EZ[q]) * uz
        f[q] = W32[q] * rho * (f32(1.0) + f32(3.0) * eu + f32(4.5) * eu * eu - f32(1.5) * u2)
    return f


def poiseuille_channel(nx: int, ny: int, nz: int, fx_lu: float, nu_lu: float, rho0: float = 1.0):
    flags = np.full((nx, ny, nz), FLUID, np.uint8)
    flags[:, 0, :] = SOLID
    flags[:, ny - 1, :] = SOLID
    flags[:, :, 0] = SOLID
    flags[:, :, nz - 1] = SOLID
    h = f32(ny - 2)
    y = (np.arange(ny, dtype=np.float32) - f32(1.0))
    prof = (f32(fx_lu) / (f32(2.0) * f32(nu_lu))) * y * (h - f32(1.0) - y)
    ux = np.zeros((nx, ny, nz), np.float32)
    ux[...] = prof[None, :, None]
    ux[flags != FLUID] = 0.0
    zero = np.zeros_like(ux)
    rho = np.full_like(ux, f32(rho0))
    f = equilibrium(rho, ux, zero, zero)
    for q in range(19):
        f[q][flags != FLUID] = W32[q] * f32(rho0)
    return dict(f=f, flags=flags, ux=ux, rho=rho, f_l=np.ones_like(ux), phi=np.ones_like(ux))


def gaussian_pulse(n: int, dx: float, T_bg: float, A: float, sigma_m: float, cp_rho: float):
    c = f32((n // 2) * dx)
    idx = np.arange(n, dtype=np.float32) * f32(dx) - c
    px, py, pz = np.meshgrid(idx, idx, idx, indexing="ij")
    r2 = px * px + py * py + pz * pz
    inv2s2 = f32(1.0) / (f32(2.0) * f32(sigma_m) * f32(sigma_m))
    return r2, inv2s2


def soft_sphere_phi(n: int, R: float, delta: float = 1.0) -> np.ndarray:
    c = (n - 1) / 2.0
    i = np.arange(n, dtype=np.float64)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    r = np.sqrt((I - c) ** 2 + (J - c) ** 2 + (K - c) ** 2)
    return (0.5 * (1.0 - np.tanh((r - R) / max(delta, 0.5)))).astype(np.float32)


def seed_pool(T, f_l, phi, flags, H, nz_solid, T_liquidus, cp_rho, L_rho, radius=5,
              center=None):
    """Seeded molten hemisphere on the plate plus a soft interface layer above it."""
    nx, ny, nz = T.shape
    i0, j0 = (nx // 3, ny // 2) if center is None else center
    k0 = nz_solid
    for di in range(-radius, radius + 1):
        for dj in range(-radius, radius + 1):
            for dk in range(0, radius):
                if di * di + dj * dj + dk * dk > radius * radius:
                    continue
                i, j, k = i0 + di, j0 + dj, k0 + dk
                if not (0 <= i < nx and 0 <= j < ny and 0 <= k < nz):
                    continue
                if flags[i, j, k] == SOLID and k < nz_solid:
                    continue
                Tm = T_liquidus + 400.0
                T[i, j, k] = Tm
                f_l[i, j, k] = 1.0
                phi[i, j, k] = 1.0
                flags[i, j, k] = FLUID
                H[i, j, k] = cp_rho * Tm + L_rho
    for di in range(-radius, radius + 1):
        for dj in range(-radius, radius + 1):
            i, j, k = i0 + di, j0 + dj, k0 + radius
            if 0 <= i < nx and 0 <= j < ny and 0 <= k < nz:
                phi[i, j, k] = 0.5
                f_l[i, j, k] = 1.0
                flags[i, j, k] = IFACE
                T[i, j, k] = T_liquidus + 200.0
                H[i, j, k] = cp_rho * T[i, j, k] + L_rho


def stefan_lambda(ste: float) -> float:
    """Root of sqrt(pi)*lam*exp(lam^2)*erf(lam) = Ste (bisection)."""
    f = lambda lam: math.sqrt(math.pi) * lam * math.exp(lam * lam) * math.erf(lam) - ste  # noqa
    lo, hi = 1e-6, 5.0
    while f(hi) < 0:
        hi *= 2.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def stefan_interface_height(fl, T, flags, nz_solid, dx, T_s, T_l):
    """Height of the (T_s+T_l)/2 isotherm above the substrate in the centre column."""
    mi, mj = fl.shape[0] // 2, fl.shape[1] // 2
    cT, cf, cfl = T[mi, mj, :], flags[mi, mj, :], fl[mi, mj, :]
    T_mid = 0.5 * (T_s + T_l)
    for k in range(nz_solid, fl.shape[2] - 1):
        if cf[k] != FLUID:
            continue
        lo, hi = float(cT[k]), float(cT[k + 1])
        if lo < T_mid <= hi:
            return (k - nz_solid + (T_mid - lo) / max(hi - lo, 1e-9)) * dx
        if lo >= T_mid > hi:
            return (k - nz_solid + 1.0 - (lo - T_mid) / max(lo - hi, 1e-9)) * dx
    for k in range(nz_solid, fl.shape[2] - 1):
        if cf[k] != FLUID:
            continue
        lo, hi = float(cfl[k]), float(cfl[k + 1])
        if lo < 0.5 <= hi:
            return (k - nz_solid + (0.5 - lo) / max(hi - lo, 1e-12)) * dx
    if cfl[nz_solid] >= 0.5:
        return 0.0
    return (fl.shape[2] - nz_solid) * dx
