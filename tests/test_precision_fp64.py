"""The precision forms against the binary64 build of the same kernels (CPU oracle).

The oracle's binary64 build (liboracle_f64.so: every fp32 variable of the kernels a double)
runs the seeded-pool bead case beside the fp32 oracle and the oracle twins with 16-bit shifted
population storage, which the GPU's fp16/bf16 paths match bit for bit (tests/test_gpu_precision.py).
One step apart the fp32 and binary64 builds agree to fp32 rounding (a wrong substitution in the
binary64 build would not); over the first ten steps, before any cell changes class, the errors
order as the storage precisions do; after 100 steps the temperature error of every form stays
within 1e-3 (the form's target), while phi, f_l and |u| carry the few interface cells that
classify differently (the seeded pool is a handful of cells across, and the fp32 run itself
moves from binary64 by |u| 4e-3 over these 100 steps)."""

from __future__ import annotations

import numpy as np
import pytest


def _twins(name="bead_mrt"):
    from oracle.oracle import Material, OracleTwin
    from tests.fixtures import LCS_BEAD_ON_PLATE
    from tests.step_cases import case_kwargs, configure, fixture_arrays

    kw = case_kwargs(name)
    out = {}
    for form in ("fp64", "fp32", "fp16", "bf16"):
        ot = OracleTwin(Material(**dict(LCS_BEAD_ON_PLATE), tables={}), mode="jacobi",
                        parallel=True, precision="fp64" if form == "fp64" else "fp32",
                        population_storage=form if form in ("fp16", "bf16") else "fp32", **kw)
        configure(ot, name)
        ot.reset()
        og = ot.g
        arrs = [og.T.astype(np.float32), og.f_l.astype(np.float32), og.phi.astype(np.float32),
                og.flags.copy(), og.H.astype(np.float32)]
        fixture_arrays(*arrs, ot.nz_solid, LCS_BEAD_ON_PLATE, ot.cp_rho, ot.L_rho)
        for n, a in zip(("T", "f_l", "phi", "flags", "H"), arrs):
            getattr(og, n)[...] = a
        out[form] = ot
    return out


def _rl2(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _err(tw, ref):
    sp = lambda g: np.sqrt(g.ux.astype(np.float64) ** 2 + g.uy ** 2 + g.uz ** 2)  # noqa: E731
    e = {k: _rl2(getattr(tw.g, k), getattr(ref.g, k)) for k in ("T", "phi", "f_l")}
    e["|u|"] = _rl2(sp(tw.g), sp(ref.g))
    return e


def test_precision_forms_against_binary64():
    from tests.step_cases import torch_x

    tw = _twins()
    ref = tw["fp64"]
    errs = {}
    for s in range(100):
        for t in tw.values():
            t.step(torch_x(t.g.dx, t.g.nx, s, t.g.dt), (t.g.ny // 2) * t.g.dx, True)
        if s == 0:
            e1 = _err(tw["fp32"], ref)
            assert e1["T"] < 1e-7 and e1["phi"] < 1e-7 and e1["|u|"] < 1e-6, e1
        if s == 9:
            e10 = {f: _err(tw[f], ref) for f in ("fp32", "fp16", "bf16")}
            assert e10["fp32"]["T"] < e10["fp16"]["T"] < e10["bf16"]["T"], e10
            assert e10["fp32"]["|u|"] < e10["fp16"]["|u|"] < e10["bf16"]["|u|"], e10
    for f in ("fp32", "fp16", "bf16"):
        errs[f] = _err(tw[f], ref)
        print(f"[fp64] {f} after 100 steps: " +
              ", ".join(f"{k} {v:.2e}" for k, v in errs[f].items()))
        assert errs[f]["T"] < 1e-3
    assert errs["fp32"]["phi"] < 1e-3 and errs["fp32"]["T"] < 1e-4
