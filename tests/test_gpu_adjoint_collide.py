"""Adjoint of the BGK collision (waam_adj_collide_srt) against the oracle's binary64 build.

Seeded populations around a moving equilibrium, body forces, liquid fractions across the
Carman-Kozeny range and a mix of FLUID, IFACE, SOLID and GAS cells. Central differences of
L = <collide(f + h df, F + h dF), lam> on the binary64 oracle (liboracle_f64.so) pin the
directional derivative <df, dL/df> + <dF, dL/dF> from the GPU adjoint, for perturbations of
the populations alone, of the forces alone and of both.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NX, NY, NZ = 12, 10, 8
N = NX * NY * NZ
E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1],
              [1, 0, -1], [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], float)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OMEGA, C_DARCY = 1.25, 50.0


def _fixture(seed=21):
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(-0.05, 0.05, (3, N))
    rho = 1.0 + rng.uniform(-0.02, 0.02, N)
    eu = E @ u0
    f = (W[:, None] * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (u0 ** 2).sum(0))
         * (1 + rng.uniform(-0.05, 0.05, (19, N))))
    F = rng.uniform(-1e-3, 1e-3, (3, N))
    f_l = rng.uniform(0.0, 1.0, N)
    flags = rng.choice(np.array([0, 4, 1, 2], np.uint8), N, p=[0.5, 0.2, 0.15, 0.15])
    lam = rng.standard_normal((19, N)).astype(np.float32)
    # fp32 inputs, as the product holds them
    return (f.astype(np.float32), F.astype(np.float32), f_l.astype(np.float32), flags, lam)


def _oracle_L(f, F, f_l, flags, lam):
    from oracle.oracle import lib as olib, ptr

    L = olib(f64=True)
    fs = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    fd = np.zeros_like(fs)
    out = [np.zeros(N) for _ in range(4)]
    Fs = [np.ascontiguousarray(F[a], dtype=np.float64) for a in range(3)]
    fl = np.ascontiguousarray(f_l, dtype=np.float64)
    L.or_collide_srt(ptr(fs), ptr(fd), ptr(out[0]), ptr(out[1]), ptr(out[2]), ptr(out[3]),
                     ptr(Fs[0]), ptr(Fs[1]), ptr(Fs[2]), ptr(fl), ptr(flags), 1.0 / OMEGA, OMEGA,
                     1.0, C_DARCY, NX, NY, NZ)
    active = (flags != 1) & (flags != 2)
    post = fd.reshape(19, N)
    return float((post[:, active] * lam[:, active].astype(np.float64)).sum())


def test_collide_srt_adjoint_against_binary64_differences():
    print("[adj-collide] srt", flush=True)
    from waam_solver_amd.adjoint import collide_srt_vjp

    f, F, f_l, flags, lam = _fixture()
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    lam_f, lam_F = collide_srt_vjp(t(f), t(F[0]), t(F[1]), t(F[2]), t(f_l), t(flags), t(lam),
                                   OMEGA, C_DARCY)
    lam_f, lam_F = lam_f.cpu().numpy(), lam_F.cpu().numpy()
    inactive = (flags == 1) | (flags == 2)
    assert not lam_f[:, inactive].any() and not lam_F[:, inactive].any()
    rng = np.random.default_rng(4)
    f64, F64 = f.astype(np.float64), F.astype(np.float64)
    for kind in ("populations", "forces", "both"):
        df = rng.standard_normal(f.shape) * 1e-2 if kind != "forces" else np.zeros(f.shape)
        dF = rng.standard_normal(F.shape) * 1e-3 if kind != "populations" else np.zeros(F.shape)
        h = 1e-4
        fd = (_oracle_L(f64 + h * df, F64 + h * dF, f_l, flags, lam)
              - _oracle_L(f64 - h * df, F64 - h * dF, f_l, flags, lam)) / (2 * h)
        adj = float((df * lam_f).sum() + (dF * lam_F).sum())
        assert abs(adj - fd) <= 1e-6 * abs(fd), (kind, adj, fd)
