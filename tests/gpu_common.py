"""Shared helpers for the GPU parity tests: seeded random states and bitwise comparison."""

from __future__ import annotations

import numpy as np

from tests.fixtures import FLUID, GAS, IFACE, SOLID, W32


def random_state(nx: int, ny: int, nz: int, seed: int = 0, nz_solid: int | None = None) -> dict:
    """A seeded, physically shaped random state: plate below, gas above, a liquid patch."""
    rng = np.random.default_rng(seed)
    s = (nx, ny, nz)
    nzs = nz // 4 if nz_solid is None else nz_solid
    k = np.arange(nz)[None, None, :]
    flags = np.where(k < nzs, SOLID, GAS).astype(np.uint8) * np.ones(s, np.uint8)
    # liquid patch and interface band sitting on the plate
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    r2 = (ii - nx / 2) ** 2 + (jj - ny / 2) ** 2
    pool = (r2[:, :, None] < (min(nx, ny) / 3) ** 2) & (k >= nzs - 2) & (k < nzs + 3)
    flags[pool] = FLUID
    band = (r2[:, :, None] < (min(nx, ny) / 3 + 2) ** 2) & (k >= nzs - 2) & (k < nzs + 5) & ~pool
    flags[band] = IFACE
    # sprinkle random flags so every branch is exercised
    sprinkle = rng.random(s) < 0.05
    flags[sprinkle] = rng.choice(np.array([FLUID, SOLID, GAS, IFACE], np.uint8), sprinkle.sum())
    phi = np.where(flags == GAS, 0.0, 1.0).astype(np.float32)
    phi[flags == IFACE] = rng.uniform(0.05, 0.95, (flags == IFACE).sum())
    noise = rng.random(s) < 0.1
    phi[noise] = rng.uniform(0.0, 1.0, noise.sum()).astype(np.float32)
    f_l = np.where(flags == FLUID, 1.0, 0.0).astype(np.float32)
    f_l[flags == IFACE] = rng.uniform(0.0, 1.0, (flags == IFACE).sum())
    part = rng.random(s) < 0.1
    f_l[part] = rng.uniform(0.0, 1.0, part.sum()).astype(np.float32)
    T = rng.uniform(300.0, 3000.0, s).astype(np.float32)
    H = (T * np.float32(5.3e6)).astype(np.float32)
    u = [(rng.standard_normal(s) * 0.03).astype(np.float32) for _ in range(3)]
    F = [(rng.standard_normal(s) * 0.01).astype(np.float32) for _ in range(3)]
    f = (W32[:, None, None, None] * (1.0 + 0.05 * rng.standard_normal((19,) + s))).astype(np.float32)
    f2 = (W32[:, None, None, None] * (1.0 + 0.05 * rng.standard_normal((19,) + s))).astype(np.float32)
    return dict(flags=flags, phi=phi, f_l=f_l, T=T, H=H, ux=u[0], uy=u[1], uz=u[2], Fx=F[0],
                Fy=F[1], Fz=F[2], f=f, f2=f2, rho=rng.uniform(0.9, 1.1, s).astype(np.float32),
                nz_solid=nzs)


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == np.float32:
        av, bv = a.view(np.uint32), b.view(np.uint32)
        same = av == bv
        both_nan = np.isnan(a) & np.isnan(b)
        return bool(np.all(same | both_nan))
    return bool(np.array_equal(a, b))


def mismatch_report(name: str, a: np.ndarray, b: np.ndarray) -> str:
    diff = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b)))) if a.dtype != np.uint8 else np.argwhere(a != b)
    if diff.size == 0:
        return f"{name}: equal"
    idx = tuple(diff[0])
    return (f"{name}: {len(diff)} mismatches, first at {idx}: gpu={a[idx]!r} oracle={b[idx]!r}")


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum(a * a)))
