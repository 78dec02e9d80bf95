"""Adjoint of the Marangoni force (waam_adj_marangoni) against the oracle's binary64 build.

The force is linear in T for held phi, f_l and flags, so the adjoint identity
<F(dT), lam> = <dT, J^T lam> holds exactly; the left side is computed by the binary64 oracle
(liboracle_f64.so) on the seeded-pool fixture, the right side by the GPU adjoint. Central
differences of L = <F(T; dgamma), lam> in dgamma on the binary64 oracle pin dL/d(dgamma), and
central differences of L in single cells of T (binary64) pin dL/dT cell by cell, including
cells on the domain faces where the stencil clamps.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.step_cases import build_pair

pytestmark = pytest.mark.gpu


def _oracle_force(T, phi, f_l, flags, dgamma):
    from oracle.oracle import lib as olib, ptr

    L = olib(f64=True)
    nx, ny, nz = T.shape
    F = [np.zeros(T.shape, dtype=np.float64) for _ in range(3)]
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (T, phi, f_l)]
    fl8 = np.ascontiguousarray(flags, dtype=np.uint8)
    L.or_compute_marangoni_force(ptr(args[0]), ptr(args[1]), ptr(args[2]), ptr(F[0]), ptr(F[1]),
                                 ptr(F[2]), ptr(fl8), float(dgamma), nx, ny, nz)
    return F


def _setup(seed=3):
    from waam_solver_amd.field import DeviceField, tensor_of

    tw, _ = build_pair("bead_mrt")
    g = tw.grid
    T = g.T.to_numpy().astype(np.float64)
    phi = g.phi.to_numpy().astype(np.float64)
    f_l = g.f_l.to_numpy().astype(np.float64)
    flags = g.flags.to_numpy().astype(np.uint8)
    rng = np.random.default_rng(seed)
    lam = [rng.standard_normal(T.shape).astype(np.float32) for _ in range(3)]
    dev = []
    for a in lam:
        f = DeviceField.zeros(T.shape, torch.float32, g.device, layout=g.layout)
        f.from_numpy(a)
        dev.append(tensor_of(f))
    return tw, g, (T, phi, f_l, flags), lam, dev


def _vjp(tw, g, lam_dev, dgamma):
    from waam_solver_amd.adjoint import marangoni_force_vjp
    from waam_solver_amd.field import DeviceField, tensor_of

    T_adj, d_dg = marangoni_force_vjp(tensor_of(g.T), tensor_of(g.phi), tensor_of(g.f_l),
                                      tensor_of(g.flags), lam_dev, dgamma, (g.nx, g.ny, g.nz),
                                      g.layout)
    return DeviceField(T_adj, layout=g.layout).to_numpy(), float(d_dg.item())


def _L(F, lam):
    return sum(float(np.dot(F[a].ravel(), lam[a].astype(np.float64).ravel())) for a in range(3))


def test_marangoni_adjoint_identity_and_dgamma():
    print("[adj-marangoni] identity", flush=True)
    tw, g, (T, phi, f_l, flags), lam, lam_dev = _setup()
    dgamma = float(tw.dgamma_dT_lu)
    active = (flags != 1) & (flags != 2) & (f_l >= 0.05)
    assert active.sum() > 50, "the fixture has no molten surface cells"
    T_adj, d_dg = _vjp(tw, g, lam_dev, dgamma)
    rng = np.random.default_rng(11)
    for _ in range(3):
        dT = rng.standard_normal(T.shape)
        lhs = _L(_oracle_force(dT, phi, f_l, flags, dgamma), lam)
        rhs = float(np.dot(dT.ravel(), T_adj.ravel()))
        assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1e-30), (lhs, rhs)
    # dL/d(dgamma) by central differences on the binary64 oracle (L is linear in dgamma)
    h = 1e-3 * abs(dgamma)
    fd = (_L(_oracle_force(T, phi, f_l, flags, dgamma + h), lam)
          - _L(_oracle_force(T, phi, f_l, flags, dgamma - h), lam)) / (2 * h)
    assert abs(d_dg - fd) <= 1e-5 * abs(fd), (d_dg, fd)


def test_marangoni_adjoint_cellwise_fd_with_clamped_faces():
    """dL/dT at single cells by central differences of the binary64 forward: cells around the
    surface and cells on the x/y faces, where the clamped stencil counts a cell twice."""
    print("[adj-marangoni] cellwise", flush=True)
    tw, g, (T, phi, f_l, flags), lam, lam_dev = _setup(seed=5)
    dgamma = float(tw.dgamma_dT_lu)
    # put molten surface cells on the x = 0 and y = ny-1 faces too
    nzs = tw.nz_solid
    f_l[0, :, nzs - 1] = 1.0
    flags[0, :, nzs - 1] = 4
    f_l[:, -1, nzs - 1] = 1.0
    flags[:, -1, nzs - 1] = 4
    g.f_l.from_numpy(f_l.astype(np.float32))
    g.flags.from_numpy(flags.astype(np.int32))
    T_adj, _ = _vjp(tw, g, lam_dev, dgamma)
    active = np.argwhere((flags != 1) & (flags != 2) & (f_l >= 0.05))
    picks = [tuple(active[i]) for i in np.linspace(0, len(active) - 1, 12).astype(int)]
    picks += [(0, 5, nzs - 1), (0, 6, nzs), (3, g.ny - 1, nzs - 1), (4, g.ny - 1, nzs - 2)]
    h = 1e-2
    for (i, j, k) in picks:
        Tp, Tm = T.copy(), T.copy()
        Tp[i, j, k] += h
        Tm[i, j, k] -= h
        fd = (_L(_oracle_force(Tp, phi, f_l, flags, dgamma), lam)
              - _L(_oracle_force(Tm, phi, f_l, flags, dgamma), lam)) / (2 * h)
        # fp32 normals in the adjoint against binary64 ones in the oracle: 1e-5 relative, plus
        # 1e-6 of the largest entry for cells where the neighbours' terms cancel
        tol = 1e-5 * abs(fd) + 1e-6 * float(np.abs(T_adj).max())
        assert abs(T_adj[i, j, k] - fd) <= tol, ((i, j, k), T_adj[i, j, k], fd)


def test_marangoni_autograd_function():
    """MarangoniForceFn: torch's backward of a loss on the force gives the adjoint's dL/dT and
    dL/d(dgamma)."""
    print("[adj-marangoni] autograd", flush=True)
    from waam_solver_amd.adjoint import MarangoniForceFn
    from waam_solver_amd.field import tensor_of

    tw, g, _, lam, lam_dev = _setup(seed=9)
    dgamma = float(tw.dgamma_dT_lu)
    T = tensor_of(g.T).clone().requires_grad_(True)
    dg = torch.tensor(dgamma, dtype=torch.float64, requires_grad=True)
    F = MarangoniForceFn.apply(T, dg, tensor_of(g.phi), tensor_of(g.f_l), tensor_of(g.flags),
                               (g.nx, g.ny, g.nz), g.layout)
    loss = sum((F[a].double() * lam_dev[a].double()).sum() for a in range(3))
    loss.backward()
    T_adj, d_dg = _vjp(tw, g, lam_dev, dgamma)
    from waam_solver_amd.field import DeviceField
    gT = DeviceField(T.grad.detach().contiguous(), layout=g.layout).to_numpy()
    assert np.allclose(gT, T_adj.astype(np.float32), rtol=0, atol=1e-6 * np.abs(T_adj).max())
    assert abs(dg.grad.item() - d_dg) <= 1e-12 * abs(d_dg)
    # the loss is linear in dgamma: loss / dgamma equals its derivative
    assert abs(loss.item() / dgamma - d_dg) <= 1e-5 * abs(d_dg)
