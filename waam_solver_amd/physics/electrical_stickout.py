"""Droplet entry temperature (reference physics/electrical_stickout.py:15-36)."""

from __future__ import annotations

import math

from .deposition_balance import wire_mass_flux_kg_s


def stickout_resistance_ohm(twin) -> float:
    L = max(twin.stickout_m, 1e-6)
    area = math.pi * (twin.wire_diameter_m * 0.5) ** 2
    return twin.rho_e_ohm_m * L / max(area, 1e-12)


def droplet_entry_temperature_K(twin) -> float:
    """Liquidus + superheat, plus I^2 R stick-out preheat when CTWD coupling is on."""
    base = twin.mat.T_liquidus + twin.deposition_superheat_K
    if not twin.enable_ctwd:
        return base
    P = twin.eta_stick * twin.welding_current_A ** 2 * stickout_resistance_ohm(twin)
    mdot = wire_mass_flux_kg_s(twin)
    if mdot <= 0:
        return base
    return min(twin.T_vapor_cap_K - 50.0, base + P / (mdot * twin.mat.cp))
