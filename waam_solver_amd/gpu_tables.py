"""Material property tables in the form the device kernels take (reference gpu_tables.py:11-66).

``MaterialGPUTables(mat)`` packs a material's piecewise-linear cp(T), k(T), mu(T) and
dgamma/dT(T) knots into the ABI's ``waam_material_tables`` struct (at most 8 knots each; an
empty table becomes the single knot (300 K, constant), as in the reference), which the
library's refresh kernel receives by value.
"""

from __future__ import annotations

import ctypes

from .materials import MaterialProps

MAX_KNOTS = 8
_F8 = ctypes.c_float * MAX_KNOTS


class MaterialTablesC(ctypes.Structure):
    """include/waam_hip.h waam_material_tables."""

    _fields_ = [("cp_T", _F8), ("cp_V", _F8), ("k_T", _F8), ("k_V", _F8), ("mu_T", _F8),
                ("mu_V", _F8), ("dgamma_T", _F8), ("dgamma_V", _F8), ("n_cp", ctypes.c_int32),
                ("n_k", ctypes.c_int32), ("n_mu", ctypes.c_int32), ("n_dgamma", ctypes.c_int32),
                ("cp_fallback", ctypes.c_float), ("k_fallback", ctypes.c_float),
                ("mu_fallback", ctypes.c_float), ("dgamma_fallback", ctypes.c_float)]


class PropertyRefsC(ctypes.Structure):
    """include/waam_hip.h waam_property_refs."""

    _fields_ = [(n, ctypes.c_float) for n in ("rho", "dt", "dx", "force_scale", "cp_rho_ref",
                                              "alpha_lu_ref", "dgamma_lu_ref", "tau_ref",
                                              "marangoni_scale")]


def pack_knots(knots, fallback: float) -> tuple[list[float], list[float], int]:
    """(t[8], v[8], n) of a knot list, reference gpu_tables.py:39-58."""
    n = min(len(knots), MAX_KNOTS)
    t = [0.0] * MAX_KNOTS
    v = [0.0] * MAX_KNOTS
    if n == 0:
        t[0], v[0], n = 300.0, float(fallback), 1
    else:
        for i in range(n):
            t[i], v[i] = float(knots[i][0]), float(knots[i][1])
    return t, v, n


class MaterialGPUTables:
    """Piecewise-linear material knots for the device kernels (max 8 points per property)."""

    def __init__(self, mat: MaterialProps):
        tbl = mat.tables
        self.enabled = bool(tbl.cp or tbl.k or tbl.mu or tbl.dgamma_dT)
        self.cp_fallback = float(mat.cp)
        self.k_fallback = float(mat.k)
        self.mu_fallback = float(mat.mu)
        self.dgamma_fallback = float(mat.dgamma_dT)
        s = MaterialTablesC()
        for prop, knots, fb in (("cp", tbl.cp, mat.cp), ("k", tbl.k, mat.k), ("mu", tbl.mu, mat.mu),
                                ("dgamma", tbl.dgamma_dT, mat.dgamma_dT)):
            t, v, n = pack_knots(knots, fb)
            getattr(s, f"{prop}_T")[:] = t
            getattr(s, f"{prop}_V")[:] = v
            setattr(s, f"n_{prop}", n)
        s.cp_fallback, s.k_fallback = self.cp_fallback, self.k_fallback
        s.mu_fallback, s.dgamma_fallback = self.mu_fallback, self.dgamma_fallback
        self.struct = s

    def arrays(self) -> tuple:
        """(t, v, n) triples in the order cp, k, mu, dgamma (for the oracle and tests)."""
        s = self.struct
        return tuple((list(getattr(s, f"{p}_T")), list(getattr(s, f"{p}_V")), getattr(s, f"n_{p}"))
                     for p in ("cp", "k", "mu", "dgamma"))


__all__ = ["MAX_KNOTS", "MaterialGPUTables", "MaterialTablesC", "PropertyRefsC", "pack_knots"]
