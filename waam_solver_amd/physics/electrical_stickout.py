"""Droplet entry temperature and the open-loop CTWD update (reference
physics/electrical_stickout.py:15-59)."""

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


def measure_bead_height_m(twin, g) -> float:
    """Height of the deposited SOLID above the plate, (top k - nz_solid + 1) dx, from a device
    reduction and one 4-byte read (reference physics/electrical_stickout.py:39-50)."""
    from .. import kernels

    kernels.bead_top_k(g.flags, twin.nz_solid, g.bead_top_buf)
    top = int(g.bead_top_buf.to_numpy()[0])
    if top < 0:
        return 0.0
    return float((top - twin.nz_solid + 1) * g.dx)


def update_ctwd(twin, g) -> None:
    """CTWD from the measured bead height against the nominal layer step (reference
    physics/electrical_stickout.py:53-59)."""
    if not twin.enable_ctwd:
        return
    h_meas = measure_bead_height_m(twin, g)
    twin._bead_height_m = h_meas
    twin.ctwd_m = max(twin.ctwd_nominal_m, twin.ctwd_nominal_m + (h_meas - twin.layer_height_m))
