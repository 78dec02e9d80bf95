"""Adjoint of the stream (waam_adj_stream) against the oracle's binary64 stream.

The stream is linear, so the adjoint identity <S df, lam> = <df, S^T lam> holds to rounding;
the left side is the binary64 oracle's stream (liboracle_f64.so) of a random df, the right
side the GPU adjoint of a random lam, on a grid whose flags mix FLUID, IFACE, SOLID and GAS
cells, including SOLID and GAS cells on the clamped y and z faces and across the periodic x
seam. A second case takes the columns of S themselves: the adjoint of a unit lam at one
output is the input that output copies. Both storage orders: the reference's (x slowest) and
the twin's default (x fastest), for which the oracle's arrays are transposed in and out.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NX, NY, NZ = 7, 6, 5
N = NX * NY * NZ


def _flags(seed=2):
    rng = np.random.default_rng(seed)
    fl = rng.choice(np.array([0, 4, 1, 2], np.uint8), (NX, NY, NZ), p=[0.4, 0.2, 0.2, 0.2])
    fl[:, 0, :] = 1      # a SOLID face
    fl[:, :, -1] = 2     # a GAS face
    fl[0, 2, 2] = 1      # SOLID on the periodic seam
    return fl


def _oracle_stream(f, flags):
    from oracle.oracle import lib as olib, ptr

    L = olib(f64=True)
    src = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    dst = np.zeros_like(src)
    L.or_stream(ptr(src), ptr(dst), ptr(np.ascontiguousarray(flags)), NX, NY, NZ)
    return dst.reshape(19, N)


def _adjoint(lam, flags, layout):
    """The GPU adjoint of lam (19, N in the oracle's x-slowest order) in the given storage
    order, returned in the oracle's order."""
    from waam_solver_amd._lib import LAYOUT_ZYX, lib

    lam4 = np.asarray(lam, dtype=np.float64).reshape(19, NX, NY, NZ)
    fl3 = np.asarray(flags).reshape(NX, NY, NZ)
    if layout == LAYOUT_ZYX:
        lam4 = lam4.transpose(0, 3, 2, 1)
        fl3 = fl3.transpose(2, 1, 0)
    dev = torch.device("cuda:0")
    lo = torch.from_numpy(np.ascontiguousarray(lam4)).to(dev)
    li = torch.empty_like(lo)
    fl = torch.from_numpy(np.ascontiguousarray(fl3)).to(dev)
    lib().call("waam_adj_stream", lo.data_ptr(), li.data_ptr(), fl.data_ptr(), NX, NY, NZ,
               layout, torch.cuda.current_stream().cuda_stream)
    out = li.cpu().numpy()
    if layout == LAYOUT_ZYX:
        out = out.transpose(0, 3, 2, 1)
    return np.ascontiguousarray(out).reshape(19, N)


@pytest.mark.parametrize("layout", [0, 1])
def test_stream_adjoint_identity(layout):
    print(f"[adj-stream] identity, layout {layout}", flush=True)
    flags = _flags()
    rng = np.random.default_rng(8)
    lam = rng.standard_normal((19, N))
    adj = _adjoint(lam, flags, layout)
    solid = (flags.reshape(-1) == 1)
    for _ in range(3):
        df = rng.standard_normal((19, N))
        out = _oracle_stream(df, flags)
        out[:, solid] = 0.0  # the stream leaves SOLID outputs unwritten (zero here)
        lhs = float((out * lam).sum())
        rhs = float((df * adj).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (lhs, rhs)


@pytest.mark.parametrize("layout", [0, 1])
def test_stream_adjoint_unit_outputs(layout):
    print(f"[adj-stream] unit outputs, layout {layout}", flush=True)
    flags = _flags(seed=5)
    rng = np.random.default_rng(1)
    fin = rng.standard_normal((19, N))
    ref = _oracle_stream(fin, flags)
    flat = flags.reshape(-1)
    for _ in range(40):
        q, c = int(rng.integers(19)), int(rng.integers(N))
        lam = np.zeros((19, N))
        lam[q, c] = 1.0
        adj = _adjoint(lam, flags, layout)
        # the adjoint of a unit output selects the input that output copies
        expect = 0.0 if flat[c] == 1 else ref[q, c]
        assert abs(float((adj * fin).sum()) - expect) <= 1e-12, (q, c, flat[c])
        assert np.count_nonzero(adj) == (0 if flat[c] == 1 else 1)
