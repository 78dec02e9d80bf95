"""ctypes binding of libwaam_hip.so (the C ABI declared in include/waam_hip.h).

The library is built in-tree by ``waam_solver_amd/build.py`` (``__graft_entry__.build()``).
Nothing here falls back to another implementation: if the library is missing or a call fails,
the error is raised.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

# torch first: the library then binds to the HIP runtime torch has loaded, so device memory
# and streams from torch are valid in it (one runtime per process)
import torch  # noqa: F401

LIB_PATH = Path(__file__).resolve().parent / "libwaam_hip.so"

ABI_VERSION = 5
FLAG_FLUID, FLAG_SOLID, FLAG_GAS, FLAG_IFACE = 0, 1, 2, 4
LAYOUT_XYZ, LAYOUT_ZYX = 0, 1

_T = {
    "p": ctypes.c_void_p,
    "f": ctypes.c_float,
    "d": ctypes.c_double,
    "i": ctypes.c_int,
    "l": ctypes.c_int64,
    "u": ctypes.c_uint32,
}

# Argument codes per entry point, in declaration order (include/waam_hip.h; tests/test_abi.py
# derives them from the header and compares): p pointer, f float, d double, i int, l int64,
# u uint32. The trailing "p" of every compute entry is the hipStream_t; grid sizes nx, ny, nz
# are followed by the field layout.
SIGNATURES: dict[str, str] = {
    "waam_init_grid": "ppppppppppfffiiiiiiiiip",
    "waam_inject_arc_heat": "ppppfffffffffiiiiipp",
    "waam_inject_goldak_heat": "ppppfffffffffffffffiiiiipp",
    "waam_surface_height_at": "pppiiiiiiip",
    "waam_clamp_enthalpy_ceiling_scalar": "ppffffflp",
    "waam_clamp_enthalpy_floor_scalar": "pfpflp",
    "waam_advect_diffuse_temperature": "pppppppfffiiiip",
    "waam_apply_thermal_boundary_losses": "pppfffiiffffiiiip",
    "waam_update_phase": "pppfffflp",
    "waam_update_T_max": "ppplp",
    "waam_update_cooling_rate": "ppppflp",
    "waam_update_time_above_T": "pppppfffflp",
    "waam_advect_phi": "ppppppiiiip",
    "waam_reinitialize_phi": "pplp",
    "waam_apply_contact_angle_phi_bc": "pppffiiiip",
    "waam_update_flags_from_phi": "pppiiiiip",
    "waam_sync_phi_liquid_fraction": "ppplp",
    "waam_solidify_cooled_metal": "pppppppfilp",
    "waam_remelt_hot_solid_scalar": "pppppfffflp",
    "waam_solidify_trailing_pool_scalar": "ppppppppffffffffffiiiip",
    "waam_compute_csf_tension": "pppppfiffiiiip",
    "waam_compute_curvature_field": "pppiiiip",
    "waam_compute_marangoni_force": "pppppppfiiiip",
    "waam_add_buoyancy": "ppppfffflp",
    "waam_add_hydrostatic_gravity": "pppfflp",
    "waam_apply_arc_pressure": "pppffffffffiiiip",
    "waam_clamp_body_force_magnitude": "ppppflp",
    "waam_apply_droplet_impact_pressure": "ppppffffffffiiiip",
    "waam_collide_mrt": "pppppppppppfffiiiip",
    "waam_collide_srt": "pppppppppppffffiiiip",
    "waam_collide_srt_variable_tau": "ppppppppppppfiiiip",
    "waam_stream": "pppiiiip",
    "waam_clamp_velocity_mach": "ppppflp",
    "waam_feed_wire_surface": "pppppppffffffffffppfiiiipp",
    "waam_deposit_droplet": "pppppppfffdfdffffppfpiiiipp",
    "waam_feed_wire_momentum_impact": "pppppppfffffffiiiip",
    "waam_inject_tracers": "pppiffffiup",
    "waam_advect_tracers": "pppppppffiiiiip",
    "waam_telemetry_pool_reduce": "ppppppppiiiipp",
    "waam_step_fused": "ppp",
    "waam_step_fused_droplet": "pppp",
    "waam_step_fused_noalloc": "ppp",
    "waam_stream_pending": "pppiiiiip",
    "waam_pack_populations": "pplip",
    "waam_unpack_populations": "pplip",
    "waam_step_arc_readback": "pp",
    "waam_step_droplet_arc_k": "pp",
    "waam_adj_marangoni": "pppppppfpppiiiip",
    "waam_adj_collide_srt": "pppppppffpplp",
    "waam_adj_stream": "pppiiiip",
    "waam_cadj_collide_mrt": "ppppppppddddppplp",
    "waam_cadj_forces": "pppppppddddippplp",
    "waam_cadj_marangoni": "ppppdppiiiip",
    "waam_cadj_phase": "pppppplp",
    "waam_cadj_thermal_step": "ppppppppppppppppppiiiipp",
    "waam_graph_capture": "pppip",
    "waam_graph_capture_plan": "pppippp",
    "waam_graph_launch": "pppppp",
    "waam_graph_steps": "p",
    "waam_graph_destroy": "p",
    "waam_timing_begin": "ii",
    "waam_timing_begin_every": "iii",
    "waam_timing_end": "pp",
    "waam_refresh_thermal_properties": "pppppppiplp",
    "waam_advect_diffuse_temperature_variable": "pppppppppfiiiip",
    "waam_update_phase_variable_cp": "ppppffffflp",
    "waam_clamp_enthalpy_floor": "pppflp",
    "waam_clamp_enthalpy_ceiling_variable_cp": "pppffflp",
    "waam_compute_marangoni_force_variable": "ppppppppiiiip",
    "waam_remelt_hot_solid": "ppppppffflp",
    "waam_solidify_trailing_pool": "pppppppppfffffffffiiiip",
    "waam_shift_simulation_window_x": "i" + "p" * 24 + "fffiiiiip",
    "waam_bead_top_k": "pipiiiip",
    "waam_apply_vapor_recoil": "ppppfffffffffiiiip",
    "waam_apply_vapor_recoil_clausius_clapeyron": "ppppffffffffffffiiiip",
    "waam_apply_gas_shear_stress": "ppppffffffffiiiip",
    "waam_elec_build_sigma": "pppfflp",
    "waam_elec_clear_source": "plp",
    "waam_elec_inject_arc_source": "pffffffiiiip",
    "waam_elec_normalize_source": "pfflpp",
    "waam_elec_init_ground": "plp",
    "waam_elec_jacobi_step": "pppppffiiiiip",
    "waam_elec_l1_diff": "ppplpp",
    "waam_elec_compute_J": "ppppppfiiiip",
    "waam_elec_bin_axial_current": "ppffffiiiiip",
    "waam_elec_prefix_bins": "piip",
    "waam_elec_B_axisymmetric": "ppppfffffiiiiip",
    "waam_apply_lorentz_JxB": "pppppppppppffflp",
    "waam_compute_vorticity_magnitude": "pppppffiiiip",
    "waam_max_norm3": "pppplp",
    "waam_probe_gather": "pppilpp",
    "waam_cavity_step": "pppppppfffffiiiip",
    "waam_adj_arc_sums": "ppppiiiipp",
    "waam_adj_terminal": "pppppplp",
    "waam_adj_thermal_step": "ppppppppppppiiiipp",
}

# kernel ids of waam_step_fused for waam_timing_begin (include/waam_hip.h WAAM_K_*)
STEP_KERNELS = ("k_step_begin", "k_step_thermal", "k_step_thermal_surface", "k_step_vof_flags",
                "k_step_collide", "k_advect_tracers_step", "k_step_stream", "k_step_pull")


class WaamError(RuntimeError):
    """A libwaam_hip entry point returned an error code."""


class _Library:
    def __init__(self, path: Path = LIB_PATH):
        if not path.exists():
            raise WaamError(
                f"{path.name} is not built; run __graft_entry__.build() "
                "(or python waam_solver_amd/build.py) first")
        self.dll = ctypes.CDLL(str(path))
        self.dll.waam_abi_version.restype = ctypes.c_int
        self.dll.waam_last_error.restype = ctypes.c_char_p
        self.dll.waam_scratch_bytes.restype = ctypes.c_int64
        self.dll.waam_noalloc_max_cells.restype = ctypes.c_int64
        self.dll.waam_quiet_bytes.restype = ctypes.c_int64
        self.dll.waam_quiet_bytes.argtypes = [ctypes.c_int] * 4
        ver = self.dll.waam_abi_version()
        if ver != ABI_VERSION:
            raise WaamError(f"libwaam_hip ABI {ver} != expected {ABI_VERSION}")
        self.fns = {}
        for name, sig in SIGNATURES.items():
            fn = getattr(self.dll, name)
            fn.argtypes = [_T[c] for c in sig.replace(" ", "")]
            fn.restype = ctypes.c_int
            self.fns[name] = fn

    def scratch_bytes(self) -> int:
        return int(self.dll.waam_scratch_bytes())

    def quiet_bytes(self, nx: int, ny: int, nz: int, layout: int) -> int:
        return int(self.dll.waam_quiet_bytes(nx, ny, nz, layout))

    def last_error(self) -> str:
        msg = self.dll.waam_last_error()
        return msg.decode() if msg else ""

    def call(self, name: str, *args) -> None:
        rc = self.fns[name](*args)
        if rc != 0:
            raise WaamError(f"{name} failed ({rc}): {self.last_error()}")


_LIB: _Library | None = None


def lib() -> _Library:
    global _LIB
    if _LIB is None:
        _LIB = _Library()
    return _LIB


def exported_symbols() -> list[str]:
    return ["waam_abi_version", "waam_last_error", "waam_scratch_bytes", "waam_noalloc_max_cells",
            "waam_quiet_bytes", *SIGNATURES]
