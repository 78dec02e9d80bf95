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

ABI_VERSION = 1
FLAG_FLUID, FLAG_SOLID, FLAG_GAS, FLAG_IFACE = 0, 1, 2, 4

_T = {
    "p": ctypes.c_void_p,
    "f": ctypes.c_float,
    "d": ctypes.c_double,
    "i": ctypes.c_int,
    "l": ctypes.c_int64,
    "u": ctypes.c_uint32,
}

# Argument codes per entry point, in declaration order (include/waam_hip.h); the trailing
# "p" of every compute entry is the hipStream_t.
SIGNATURES: dict[str, str] = {
    "waam_init_grid": "pppppppppp fff iiiii iii p",
    "waam_inject_arc_heat": "pppp fffffffff i iii pp",
    "waam_inject_goldak_heat": "pppp fffffffffffffff i iii pp",
    "waam_surface_height_at": "ppp iii iii p",
    "waam_clamp_enthalpy_ceiling_scalar": "pp fffff l p",
    "waam_clamp_enthalpy_floor_scalar": "pfpf l p",
    "waam_advect_diffuse_temperature": "ppppppp fff iii p",
    "waam_apply_thermal_boundary_losses": "ppp fff ii ffff iii p",
    "waam_update_phase": "ppp ffff l p",
    "waam_update_T_max": "ppp l p",
    "waam_update_cooling_rate": "pppp f l p",
    "waam_update_time_above_T": "ppppp ffff l p",
    "waam_advect_phi": "pppppp iii p",
    "waam_reinitialize_phi": "pp l p",
    "waam_apply_contact_angle_phi_bc": "ppp ff iii p",
    "waam_update_flags_from_phi": "ppp i iii p",
    "waam_sync_phi_liquid_fraction": "ppp l p",
    "waam_solidify_cooled_metal": "ppppppp f i l p",
    "waam_remelt_hot_solid_scalar": "ppppp ffff l p",
    "waam_solidify_trailing_pool_scalar": "pppppppp ffffffffff iii p",
    "waam_compute_csf_tension": "ppppp f i ff iii p",
    "waam_compute_curvature_field": "ppp iii p",
    "waam_compute_marangoni_force": "ppppppp f iii p",
    "waam_add_buoyancy": "pppp ffff l p",
    "waam_add_hydrostatic_gravity": "ppp ff l p",
    "waam_apply_arc_pressure": "ppp ffffffff iii p",
    "waam_clamp_body_force_magnitude": "pppp f l p",
    "waam_apply_droplet_impact_pressure": "pppp ffffffff iii p",
    "waam_collide_mrt": "ppppppppppp fff iii p",
    "waam_collide_srt": "ppppppppppp ffff iii p",
    "waam_collide_srt_variable_tau": "pppppppppppp f iii p",
    "waam_stream": "ppp iii p",
    "waam_clamp_velocity_mach": "pppp f l p",
    "waam_feed_wire_surface": "ppppppp ffffffffff pp f iii pp",
    "waam_deposit_droplet": "ppppppp fff d f d ffff pp f p iii pp",
    "waam_feed_wire_momentum_impact": "ppppppp fffffff iii p",
    "waam_inject_tracers": "ppp i ffff i u p",
    "waam_advect_tracers": "ppppppp ff i iii p",
    "waam_telemetry_pool_reduce": "pppppppp iii pp",
    "waam_step_fused": "ppp",
    "waam_stream_pending": "ppp iii p",
    "waam_step_arc_readback": "pp",
    "waam_timing_begin": "ii",
    "waam_timing_end": "pp",
}

# kernel ids of waam_step_fused for waam_timing_begin (include/waam_hip.h WAAM_K_*)
STEP_KERNELS = ("k_step_begin", "k_step_thermal", "k_step_thermal_surface", "k_step_vof_flags",
                "k_step_collide", "k_advect_tracers_step", "k_step_stream")


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
    return ["waam_abi_version", "waam_last_error", "waam_scratch_bytes", *SIGNATURES]
