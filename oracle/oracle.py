"""CPU oracle for the waam_twin coupled step — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``; the
product package never imports it. It wraps ``liboracle.so`` (oracle/waam_oracle.c, one C
function per reference Taichi kernel) with numpy arrays and restates, independently of the
product, the host logic around those kernels:

* lattice constants — reference grid.py:89-105 and twin.py:244-274;
* the per-step kernel order — reference solvers/coupled_step.py:73-464 (scalar-material path);
* the droplet schedule — reference physics/deposition_balance.py:12-86,
  physics/deposition.py:27-31, physics/weld_forces.py:61-119,
  physics/electrical_stickout.py:23-36;
* reset — reference twin.py:486-531; telemetry — twin.py:706-744.

``mode="jacobi"`` selects the GPU semantics for the two order-dependent in-place kernels
(thermal advection, contact-angle BC); ``mode="serial"`` selects in-place lexicographic order
for both, ``"serial_thermal"`` / ``"serial_contact"`` for one of them (the deviation of each
from the Jacobi form can then be measured on its own).
"""

from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"

FLUID, SOLID, GAS, IFACE = 0, 1, 2, 4
Q = 19
EX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
EY = (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1)
EZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1)
W = (1 / 3,) + (1 / 18,) * 6 + (1 / 36,) * 12
OPP = (0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17)

_F = ctypes.c_float
_I = ctypes.c_int
_P = ctypes.c_void_p


def build(verbose: bool = False) -> None:
    """Compile liboracle*.so with the recipe in oracle/Makefile."""
    import subprocess

    cmd = ["make", "-C", str(HERE), "-B"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(res.stdout)


def _load(name: str) -> ctypes.CDLL:
    path = BUILD / name
    if not path.exists():
        build()
    return ctypes.CDLL(str(path))


_SIGS = {
    "or_init_grid": [_P] * 10 + [_F, _F, _F, _I, _I, _I, _I, _I, _I, _I, _I],
    "or_arc_weight_sum": [_P, _P, _F, _F, _F, _F, _F, _I, _I, _I, _I],
    "or_inject_arc_heat": [_P, _P, _P, _P] + [_F] * 9 + [_I, _I, _I, _I],
    "or_inject_goldak_heat": [_P, _P, _P, _P] + [_F] * 15 + [_I, _I, _I, _I],
    "or_surface_height_at": [_P, _P, _I, _I, _I, _I, _I, _I],
    "or_feed_wire_surface": [_P] * 7 + [_F] * 10 + [_P, _P, _F, _I, _I, _I],
    "or_feed_wire_momentum_impact": [_P] * 7 + [_F] * 7 + [_I, _I, _I],
    "or_apply_droplet_impact_pressure": [_P] * 4 + [_F] * 8 + [_I, _I, _I],
    "or_compute_csf_tension": [_P] * 5 + [_F, _I, _F, _F, _I, _I, _I],
    "or_compute_curvature_field": [_P, _P, _P, _I, _I, _I],
    "or_apply_contact_angle_phi_bc_serial": [_P, _P, _F, _F, _I, _I, _I],
    "or_apply_contact_angle_phi_bc_jacobi": [_P, _P, _P, _F, _F, _I, _I, _I],
    "or_update_phase": [_P, _P, _P, _F, _F, _F, _F, _I, _I, _I],
    "or_advect_diffuse_temperature_serial": [_P] * 6 + [_F, _F, _F, _I, _I, _I],
    "or_advect_diffuse_temperature_jacobi": [_P] * 7 + [_F, _F, _F, _I, _I, _I],
    "or_apply_thermal_boundary_losses": [_P, _P, _P, _F, _F, _F, _I, _I, _F, _F, _F, _F, _I, _I, _I],
    "or_clamp_enthalpy_floor_scalar": [_P, _F, _P, _F, _I, _I, _I],
    "or_clamp_enthalpy_ceiling_scalar": [_P, _P, _F, _F, _F, _F, _F, _I, _I, _I],
    "or_update_cooling_rate": [_P, _P, _P, _P, _F, _I, _I, _I],
    "or_update_T_max": [_P, _P, _P, _I, _I, _I],
    "or_update_time_above_T": [_P] * 5 + [_F, _F, _F, _F, _I, _I, _I],
    "or_compute_marangoni_force": [_P] * 7 + [_F, _I, _I, _I],
    "or_add_buoyancy": [_P] * 4 + [_F, _F, _F, _F, _I, _I, _I],
    "or_add_hydrostatic_gravity": [_P] * 3 + [_F, _F, _I, _I, _I],
    "or_apply_arc_pressure": [_P] * 3 + [_F] * 8 + [_I, _I, _I],
    "or_clamp_body_force_magnitude": [_P] * 4 + [_F, _I, _I, _I],
    "or_advect_phi": [_P] * 6 + [_I, _I, _I],
    "or_reinitialize_phi": [_P, _P, _I, _I, _I],
    "or_update_flags_from_phi": [_P, _P, _P, _I, _I, _I, _I],
    "or_sync_phi_liquid_fraction": [_P, _P, _P, _I, _I, _I],
    "or_solidify_cooled_metal": [_P] * 7 + [_F, _I, _I, _I, _I],
    "or_remelt_hot_solid_scalar": [_P] * 5 + [_F, _F, _F, _F, _I, _I, _I],
    "or_solidify_trailing_pool_scalar": [_P] * 8 + [_F] * 10 + [_I, _I, _I],
    "or_collide_srt": [_P] * 11 + [_F, _F, _F, _F, _I, _I, _I],
    "or_collide_srt_variable_tau": [_P] * 12 + [_F, _I, _I, _I],
    "or_collide_mrt": [_P] * 11 + [_F, _F, _F, _I, _I, _I],
    "or_stream": [_P, _P, _P, _I, _I, _I],
    "or_clamp_velocity_mach": [_P] * 4 + [_F, _I, _I, _I],
    "or_telemetry_pool_reduce": [_P] * 8 + [_I, _I, _I],
    "or_inject_tracers": [_P, _P, _P, _I, _F, _F, _F, _F, _I, ctypes.c_uint32],
    "or_advect_tracers": [_P] * 7 + [_F, _F, _I, _I, _I, _I],
    "or_refresh_thermal_properties": [_P] * 8 + [_I, _P, _I, _I, _I],
    "or_advect_diffuse_temperature_variable_serial": [_P] * 8 + [_F, _I, _I, _I],
    "or_advect_diffuse_temperature_variable_jacobi": [_P] * 9 + [_F, _I, _I, _I],
    "or_update_phase_variable_cp": [_P] * 4 + [_F] * 5 + [_I, _I, _I],
    "or_clamp_enthalpy_floor": [_P, _P, _P, _F, _I, _I, _I],
    "or_clamp_enthalpy_ceiling_variable_cp": [_P, _P, _P, _F, _F, _F, _I, _I, _I],
    "or_compute_marangoni_force_variable": [_P] * 8 + [_I, _I, _I],
    "or_remelt_hot_solid": [_P] * 6 + [_F, _F, _F, _I, _I, _I],
    "or_solidify_trailing_pool": [_P] * 9 + [_F] * 9 + [_I, _I, _I],
    "or_shift_simulation_window_x": [_I, _P, _P, _P, _P, _F, _F, _F, _I, _I, _I, _I],
    "or_apply_vapor_recoil": [_P] * 4 + [_F] * 8 + [_I, _I, _I],
    "or_apply_vapor_recoil_cc": [_P] * 4 + [_F] * 11 + [_I, _I, _I],
    "or_apply_gas_shear_stress": [_P] * 4 + [_F] * 7 + [_I, _I, _I],
    "or_elec_build_sigma": [_P, _P, _P, _F, _F, _I, _I, _I],
    "or_elec_inject_arc_source": [_P] + [_F] * 6 + [_I, _I, _I],
    "or_elec_normalize_source": [_P, _F, _F, _I, _I, _I],
    "or_elec_jacobi_step": [_P] * 5 + [_F, _F, _I, _I, _I, _I],
    "or_elec_l1_diff": [_P, _P, _P, _I, _I, _I],
    "or_elec_compute_J": [_P] * 6 + [_F, _I, _I, _I],
    "or_elec_bin_axial_current": [_P, _P, _F, _F, _F, _F, _I, _I, _I, _I],
    "or_elec_prefix_bins": [_P, _I, _I],
    "or_elec_B_axisymmetric": [_P] * 4 + [_F] * 5 + [_I, _I, _I, _I],
    "or_apply_lorentz_JxB": [_P] * 11 + [_F, _F, _F, _I, _I, _I],
    "or_compute_vorticity_magnitude": [_P] * 5 + [_F, _I, _I, _I],
    "or_cavity_step": [_P] * 7 + [_F] * 5 + [_I, _I, _I],
}


class _Lib:
    def __init__(self, name: str, f64: bool = False):
        self.dll = _load(name)
        real = ctypes.c_double if f64 else _F
        for fn, sig in _SIGS.items():
            f = getattr(self.dll, fn)
            f.argtypes = [real if t is _F else t for t in sig]
            f.restype = None
        self.dll.or_arc_weight_sum.restype = real
        self.dll.or_surface_height_at.restype = real

    def __getattr__(self, item):
        return getattr(self.dll, item)


_LIBS: dict[str, _Lib] = {}


def lib(parallel: bool = False, f64: bool = False) -> _Lib:
    """The oracle library: fp32 (single-threaded, or OpenMP with ``parallel``) or the binary64
    build of the same kernels (``f64``, OpenMP; arrays float64, scalars double)."""
    name = "liboracle_f64.so" if f64 else ("liboracle_omp.so" if parallel else "liboracle.so")
    if name not in _LIBS:
        _LIBS[name] = _Lib(name, f64=f64)
    return _LIBS[name]


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


# --------------------------------------------------------------------------------------------
# Materials: the constants block of a material file plus the Heiple dγ/dT scaling
# (reference materials.py:61-82, physics/surfactant.py:28-50, :177-194).
# --------------------------------------------------------------------------------------------
@dataclass
class Material:
    name: str
    k: float
    rho: float
    cp: float
    T_solidus: float
    T_liquidus: float
    L_fusion: float
    gamma_0: float
    dgamma_dT: float
    mu: float
    beta_T: float = 1.2e-4
    alpha: float | None = None
    contact_angle_deg: float = 80.0
    sulphur_ppm: float = 30.0
    rho_e_ohm_m: float = 1.5e-7
    eta_stick: float = 0.85
    # piecewise-linear knots {"cp", "k", "mu", "dgamma_dT"} -> [(T, value), ...], given as
    # data (already through the surfactant model); empty means the scalar path
    tables: dict | None = None

    def __post_init__(self):
        if self.alpha is None:
            self.alpha = self.k / (self.rho * self.cp)
        s = max(0.0, self.sulphur_ppm)
        if s <= 30.0:
            scale = 1.0
        elif s >= 60.0:
            scale = -0.35
        else:
            scale = 1.0 - 1.35 * (s - 30.0) / 30.0
        self.dgamma_dT = -abs(self.dgamma_dT) * scale if scale < 0 else self.dgamma_dT * scale
        self.tables = {k: [(float(t), float(v)) for t, v in (self.tables or {}).get(k, [])]
                       for k in ("cp", "k", "mu", "dgamma_dT")}

    def has_tables(self) -> bool:
        return any(self.tables.values())


def interp_knots(T: float, knots, fallback: float) -> float:
    """Host piecewise-linear lookup in double precision (reference material_tables.py:42-55):
    clamped to the end values, first bracketing interval wins, zero-width interval -> left."""
    if not knots:
        return fallback
    if T <= knots[0][0]:
        return knots[0][1]
    if T >= knots[-1][0]:
        return knots[-1][1]
    for (t0, v0), (t1, v1) in zip(knots[:-1], knots[1:]):
        if t0 <= T <= t1:
            return v0 if t1 == t0 else v0 + (T - t0) / (t1 - t0) * (v1 - v0)
    return fallback


def pack_tables(mat: "Material") -> tuple[np.ndarray, np.ndarray]:
    """(64 float32 knots, 4 int32 counts) in the order cp, k, mu, dgamma: at most 8 knots per
    property, an empty property becomes the single knot (300 K, scalar) (reference
    gpu_tables.py:39-58)."""
    tab = np.zeros(64, np.float32)
    cnt = np.zeros(4, np.int32)
    for p, (key, fb) in enumerate((("cp", mat.cp), ("k", mat.k), ("mu", mat.mu),
                                   ("dgamma_dT", mat.dgamma_dT))):
        knots = mat.tables[key][:8] or [(300.0, fb)]
        for i, (t, v) in enumerate(knots):
            tab[16 * p + i] = t
            tab[16 * p + 8 + i] = v
        cnt[p] = len(knots)
    return tab, cnt


# --------------------------------------------------------------------------------------------
# Grid: numpy SoA arrays (reference grid.py:132-206). f_b starts at zero like a fresh Taichi
# field; f_a is written by init_grid.
# --------------------------------------------------------------------------------------------
class OracleGrid:
    def __init__(self, nx: int, ny: int, nz: int, dx: float, mat: Material, max_tracers: int,
                 dtype=np.float32):
        self.dtype = dtype
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.u_ref_phys, self.u_ref_lu = 0.5, 0.05
        self.dt = dx * self.u_ref_lu / self.u_ref_phys
        nu_phys = mat.mu / mat.rho
        nu_lu = nu_phys * self.dt / (dx ** 2)
        self.tau = 3.0 * nu_lu + 0.5
        self.alpha_lu = mat.alpha * self.dt / (dx ** 2)
        if self.tau <= 0.5:
            raise ValueError(f"LBM unstable: tau={self.tau:.4f} <= 0.5")
        if self.alpha_lu > 1.0 / 6.0:
            raise ValueError(f"Thermal step unstable: alpha_lu={self.alpha_lu:.4f} > 1/6")
        s3 = (nx, ny, nz)
        z = lambda: np.zeros(s3, dtype)  # noqa: E731
        self.f_a = np.zeros((Q,) + s3, dtype)
        self.f_b = np.zeros((Q,) + s3, dtype)
        for name in ("rho", "T", "T_max", "H", "f_l", "phi", "dT_dt", "T_prev",
                     "time_above_800_s", "time_above_1100_s", "time_above_solidus_s",
                     "Fx_snap", "Fy_snap", "Fz_snap", "ux", "uy", "uz", "Fx", "Fy", "Fz",
                     "phi_tmp", "cp_rho_field", "alpha_lu_field", "dgamma_lu_field",
                     "tau_field"):
            setattr(self, name, z())
        self.flags = np.zeros(s3, np.uint8)
        self.max_tracers = max_tracers
        self.porosity_pos = np.zeros((max_tracers, 3), dtype)
        self.porosity_active = np.zeros(max_tracers, np.int32)
        self.tracer_head = np.zeros(1, np.int32)
        self.deposit_vol = np.zeros(1, dtype)
        self.deposit_real = np.zeros(1, dtype)
        self.read_a = True

    @property
    def f_src(self) -> np.ndarray:
        return self.f_a if self.read_a else self.f_b

    @property
    def f_dst(self) -> np.ndarray:
        return self.f_b if self.read_a else self.f_a

    def swap_buffers(self) -> None:
        self.read_a = not self.read_a

    @property
    def dims(self) -> tuple[int, int, int]:
        return self.nx, self.ny, self.nz


class OracleTwin:
    """Independent restatement of WAAMTwin's constructor, reset and coupled step (scalar and
    temperature-dependent table paths)."""

    def __init__(self, material: Material, nx=256, ny=128, nz=64, dx=2e-4, arc_power_W=2800.0,
                 arc_efficiency=0.8, T_ambient=300.0, use_srt=True, C_darcy=1.6e5,
                 arc_pressure_pa=500.0, droplet_freq_hz=50.0, max_tracers=None,
                 enable_heat_loss=False, h_conv=25.0, eps_rad=0.3, enable_convection=True,
                 enable_radiation=False, heat_source="gaussian2d", enable_vof=False,
                 enable_csf_tension=False, enable_deposition_momentum=True,
                 enable_droplet_impact_pressure=True, enable_substrate_growth=False,
                 enable_wetting=False, enable_hydrostatic_gravity=False,
                 enable_bead_freeze=False, contact_angle_deg=None,
                 deposition_superheat_K=500.0, deposition_footprint_sigma_scale=1.0,
                 welding_current_A=140.0, travel_speed_m_s=0.0, enable_enthalpy_cap=True,
                 T_vapor_cap_K=3200.0, arc_surface_weighting=True, arc_penetration_mm=2.0,
                 wire_diameter_mm=1.2, arc_sigma_mm=2.0, bulk_tau=None, mode="jacobi",
                 parallel=False, tracer_seed=0, use_variable_tau=True,
                 enable_moving_window=False, enable_ctwd=False, enable_recoil=False,
                 use_recoil_clausius_clapeyron=True, recoil_pressure_pa=5000.0,
                 enable_gas_shear=False, gas_jet_velocity_m_s=12.0, gas_shear_coeff=1.0,
                 T_boiling_K=3100.0, P_vapor_ref_Pa=101325.0, L_vapor_J_kg=6.0e6,
                 R_spec_vapor_J_kgK=450.0, enable_lorentz=False, sigma_liquid_Sm=1.0e6,
                 sigma_solid_Sm=1.0e5, lorentz_jacobi_iters=40, lorentz_jacobi_omega=0.85,
                 population_storage="fp32", precision="fp32"):
        self.mat = material
        if precision not in ("fp32", "fp64"):
            raise ValueError(f"precision {precision!r}")
        self.precision = precision
        self.lib = lib(parallel, f64=precision == "fp64")
        self.mode = mode
        if population_storage not in ("fp32", "fp16", "bf16"):
            raise ValueError(f"population_storage {population_storage!r}")
        self.population_storage = population_storage
        self.Q_w, self.eta, self.T_amb = arc_power_W, arc_efficiency, T_ambient
        self.use_srt, self.C_darcy, self.arc_pressure = use_srt, C_darcy, arc_pressure_pa
        self.droplet_freq = droplet_freq_hz
        self.enable_heat_loss, self.h_conv, self.eps_rad = enable_heat_loss, h_conv, eps_rad
        self.enable_convection, self.enable_radiation = enable_convection, enable_radiation
        self.sigma_sb = 5.670374419e-8
        self.heat_source_name = heat_source
        self.enable_vof, self.enable_csf_tension = enable_vof, enable_csf_tension
        self.enable_deposition_momentum = enable_deposition_momentum
        self.enable_droplet_impact_pressure = enable_droplet_impact_pressure
        self.enable_substrate_growth, self.enable_wetting = enable_substrate_growth, enable_wetting
        self.enable_hydrostatic_gravity = enable_hydrostatic_gravity
        self.enable_bead_freeze = enable_bead_freeze
        self.contact_angle_deg = (contact_angle_deg if contact_angle_deg is not None
                                  else material.contact_angle_deg)
        self.theta_rad = math.radians(self.contact_angle_deg)
        self.deposition_superheat_K = deposition_superheat_K
        self.deposition_footprint_sigma_scale = deposition_footprint_sigma_scale
        self.welding_current_A, self.travel_speed_m_s = welding_current_A, travel_speed_m_s
        self.enable_enthalpy_cap, self.T_vapor_cap_K = enable_enthalpy_cap, T_vapor_cap_K
        self.arc_surface_weighting = arc_surface_weighting
        self.arc_penetration_m = arc_penetration_mm * 1e-3
        self.wire_diameter_m = wire_diameter_mm * 1e-3
        self.wire_feed_m_s = 0.0
        self.droplet_transfer_mode = "auto"
        self.pulse_frequency_hz = 0.0
        self.droplet_size_jitter = 0.12
        self.impact_lead_angle_deg = 0.0
        self.trailing_solidify_lookback_mm = 2.5
        self.trailing_solidify_temp_margin_K = 35.0
        self.u_mach_limit_lu, self.force_limit_lu = 0.08, 0.05
        self.arc_pressure_model = "constant"
        self.pressure_sigma_m = None
        self.enable_ctwd = enable_ctwd
        # full-tier surface forces (twin.py:181-197; physics/weld_forces.py:120-161)
        self.enable_recoil, self.use_recoil_cc = enable_recoil, use_recoil_clausius_clapeyron
        self.recoil_pressure_pa, self.recoil_accommodation = recoil_pressure_pa, 0.54
        self.T_boiling_K, self.P_vapor_ref_Pa = T_boiling_K, P_vapor_ref_Pa
        self.L_vapor_J_kg, self.R_spec_vapor_J_kgK = L_vapor_J_kg, R_spec_vapor_J_kgK
        self.enable_gas_shear = enable_gas_shear
        self.enable_lorentz = enable_lorentz
        self.sigma_liquid_Sm, self.sigma_solid_Sm = sigma_liquid_Sm, sigma_solid_Sm
        self.lorentz_jacobi_iters, self.lorentz_jacobi_omega = lorentz_jacobi_iters, lorentz_jacobi_omega
        self.lorentz_jacobi_tol = 1e-4
        self._lorentz_warm = False
        self._lorentz_unconverged = 0
        self.gas_jet_velocity_m_s, self.gas_shear_coeff = gas_jet_velocity_m_s, gas_shear_coeff
        # stick-out and CTWD (twin.py:186-191; physics/electrical_stickout.py:15-59)
        self.stickout_m, self.ctwd_nominal_m, self.layer_height_m = 12e-3, 15e-3, 1.2e-3
        self.ctwd_m = self.ctwd_nominal_m
        self.rho_e_ohm_m, self.eta_stick = material.rho_e_ohm_m, material.eta_stick
        self._bead_height_m = 0.0
        self.plate_thickness_mm = None
        self.plate_size_mm = None
        self.plate_origin_mm = None
        self.nz_solid = max(1, nz // 5)
        self.enable_moving_window = enable_moving_window
        self._window_offset_x_m = 0.0
        self.tracer_seed = tracer_seed
        self.g = OracleGrid(nx, ny, nz, dx, material, 50_000 if max_tracers is None else max_tracers,
                            dtype=np.float64 if precision == "fp64" else np.float32)
        g = self.g
        self.cp_rho = material.rho * material.cp
        self.L_rho = material.rho * material.L_fusion
        self.alpha_lu = material.alpha * g.dt / (g.dx ** 2)
        self.omega = 1.0 / g.tau
        self.omega_bulk = 1.0 / bulk_tau if bulk_tau else self.omega
        self.arc_sigma_m = arc_sigma_mm * 1e-3
        self.sigma_cells = self.arc_sigma_m / g.dx
        self.g_lu = 9.81 * (g.dt ** 2) / g.dx
        self.beta_T = material.beta_T
        scale_F = (g.dt ** 2) / (material.rho * g.dx ** 2)
        self.force_scale = scale_F
        self.dgamma_dT_lu = material.dgamma_dT * scale_F / g.dx
        self.gamma_lu = material.gamma_0 * scale_F / g.dx
        self.marangoni_scale = 1.0
        # temperature-dependent tables (twin.py:269-273)
        self.use_tables = material.has_tables()
        self.use_variable_tau = use_variable_tau
        self.table_knots, self.table_counts = pack_tables(material)
        cp_sol = interp_knots(material.T_solidus, material.tables["cp"], material.cp)
        self.H_sol = material.rho * cp_sol * material.T_solidus
        self.H_liq = self.H_sol + self.L_rho
        self.goldak = None  # (ff, fr, a_front, a_rear, b, c) in cells when heat_source is Goldak
        self._reset_counters()

    # ---------------------------------------------------------------- geometry (twin.py:289-342)
    def resolve_nz_solid(self, test_fluid_domain=False):
        if test_fluid_domain:
            return 0
        g = self.g
        headroom = max(6, g.nz // 4)
        max_solid = max(1, g.nz - headroom)
        if self.plate_thickness_mm is not None and float(self.plate_thickness_mm) > 0.0:
            n = int(round(float(self.plate_thickness_mm) / (g.dx * 1000.0)))
            return max(1, min(n, max_solid))
        return max(1, min(g.nz // 5, max_solid))

    def resolve_plate_ij(self):
        g = self.g
        dx_mm = g.dx * 1000.0
        if not self.plate_size_mm:
            return 0, g.nx, 0, g.ny
        lx, ly = float(self.plate_size_mm[0]), float(self.plate_size_mm[1])
        ni = min(max(1, int(round(lx / dx_mm))), g.nx)
        nj = min(max(1, int(round(ly / dx_mm))), g.ny)
        if self.plate_origin_mm is not None:
            i0 = int(round(float(self.plate_origin_mm[0]) / dx_mm))
            j0 = int(round(float(self.plate_origin_mm[1]) / dx_mm))
        else:
            i0, j0 = (g.nx - ni) // 2, (g.ny - nj) // 2
        i0 = max(0, min(i0, g.nx - ni))
        j0 = max(0, min(j0, g.ny - nj))
        return i0, i0 + ni, j0, j0 + nj

    def _reset_counters(self):
        self._step_n = 0
        self._last_droplet_time = 0.0
        self._deposited_volume_m3 = 0.0
        self._n_droplets_fired = 0
        self._deposition_overflow = 0
        self._last_torch_pos_m = None
        self._torch_dir_xyz = (1.0, 0.0, 0.0)

    # ---------------------------------------------------------------- reset (twin.py:486-531)
    def reset(self, T_ambient=None, test_fluid_domain=False):
        T0 = T_ambient or self.T_amb
        g, L = self.g, self.lib
        if test_fluid_domain:
            nzs = -1
            self.nz_solid = 0
            i0, i1, j0, j1 = 0, g.nx, 0, g.ny
        else:
            self.nz_solid = self.resolve_nz_solid(False)
            nzs = self.nz_solid
            i0, i1, j0, j1 = self.resolve_plate_ij()
        L.or_init_grid(ptr(g.f_a), ptr(g.rho), ptr(g.ux), ptr(g.uy), ptr(g.uz), ptr(g.T),
                       ptr(g.H), ptr(g.f_l), ptr(g.phi), ptr(g.flags), T0, 1.0, self.cp_rho,
                       nzs, i0, i1, j0, j1, *g.dims)
        for a in (g.T_max, g.T_prev):
            a.fill(g.dtype(T0))
        for a in (g.dT_dt, g.time_above_800_s, g.time_above_1100_s, g.time_above_solidus_s,
                  g.Fx, g.Fy, g.Fz, g.Fx_snap, g.Fy_snap, g.Fz_snap):
            a.fill(0)
        g.porosity_active.fill(0)
        g.tracer_head[0] = 0
        L.or_stream(ptr(g.f_a), ptr(g.f_b), ptr(g.flags), *g.dims)
        self._reset_counters()
        self._lorentz_warm = False
        g.deposit_vol[0] = 0
        g.deposit_real[0] = 0

    # ---------------------------------------------------------------- host helpers
    def wire_mass_flux_kg_s(self):
        d = self.wire_diameter_m
        return self.mat.rho * (math.pi * (d * 0.5) ** 2) * self.wire_feed_m_s

    def transfer_mode(self):
        mode = str(self.droplet_transfer_mode).strip().lower()
        if mode in {"globular", "spray", "pulsed"}:
            return mode
        if self.pulse_frequency_hz > 0.0:
            return "pulsed"
        return "globular" if float(self.welding_current_A) < 180.0 else "spray"

    def droplet_period_s(self):
        f = float(self.droplet_freq)
        base = 1.0 / max(f, 1e-9) if f > 0.0 else 1.0 / 50.0
        mode = self.transfer_mode()
        if mode == "globular":
            ms = 1.55
        elif mode == "spray":
            ms = 0.72
        elif mode == "pulsed":
            ms = max(0.25, min(4.0, (1.0 / max(self.pulse_frequency_hz, 1e-9)) / base))
        else:
            ms = 1.0
        idx = float(self._n_droplets_fired + 1)
        amp = max(0.0, min(float(self.droplet_size_jitter), 0.35))
        jitter = 1.0 + amp * math.sin(0.73 * idx + 0.31)
        return max(base * 0.20, base * ms * jitter)

    def droplet_mass_kg(self):
        if self.droplet_freq <= 0:
            return 0.0
        return max(0.0, self.wire_mass_flux_kg_s() * max(self.droplet_period_s(), 0.0))

    def droplet_radius_cells(self, m):
        if m <= 0:
            return 3.0
        n_cells = (m / self.mat.rho) / (self.g.dx ** 3)
        r = (3.0 * n_cells / (4.0 * math.pi)) ** (1.0 / 3.0)
        return max(1.0, min(r, 10.0))

    def arc_pressure_sigma_m(self):
        """Pressure Gaussian radius: its own when set, else the arc heat sigma
        (reference physics/weld_forces.py:35-40)."""
        if self.pressure_sigma_m is not None and self.pressure_sigma_m > 0.0:
            return float(self.pressure_sigma_m)
        return float(self.arc_sigma_m)

    def arc_pressure_peak_pa(self):
        """Lin & Eagar (1986) p0 = mu0 I^2 / (4 pi^2 sigma_p^2) for the "lin_eagar" model, else
        the constant (reference physics/weld_forces.py:20-32, :47-58)."""
        if str(self.arc_pressure_model).strip().lower() in ("lin_eagar", "lineagar", "lin-eagar"):
            sig = max(self.arc_pressure_sigma_m(), 1.0e-6)
            i = max(float(self.welding_current_A), 0.0)
            return 4.0e-7 * math.pi * i * i / (4.0 * math.pi * math.pi * sig * sig)
        return float(self.arc_pressure)

    def droplet_entry_temperature_K(self):
        """Liquidus + superheat, plus I^2 R heating of the stick-out (rho_e L / A) spread over
        the wire mass flux when CTWD coupling is on, capped 50 K under the vapour cap."""
        base = self.mat.T_liquidus + self.deposition_superheat_K
        if not self.enable_ctwd:
            return base
        area = math.pi * (self.wire_diameter_m * 0.5) ** 2
        R = self.rho_e_ohm_m * max(self.stickout_m, 1e-6) / max(area, 1e-12)
        P = self.eta_stick * self.welding_current_A ** 2 * R
        mdot = self.wire_mass_flux_kg_s()
        if mdot <= 0:
            return base
        return min(self.T_vapor_cap_K - 50.0, base + P / (mdot * self.mat.cp))

    def bead_height_m(self) -> float:
        """(top SOLID k at or above nz_solid - nz_solid + 1) dx, 0 without one (reference
        physics/electrical_stickout.py:39-50)."""
        g = self.g
        ks = np.nonzero(g.flags == SOLID)[2]
        ks = ks[ks >= self.nz_solid]
        return 0.0 if ks.size == 0 else float((ks.max() - self.nz_solid + 1) * g.dx)

    def update_ctwd(self):
        """CTWD follows the bead height against the nominal layer height, never below nominal
        (reference physics/electrical_stickout.py:53-59)."""
        h = self.bead_height_m()
        self._bead_height_m = h
        self.ctwd_m = max(self.ctwd_nominal_m, self.ctwd_nominal_m + (h - self.layer_height_m))

    def toe_angle_deg(self) -> float:
        """Reference physics/bead_geometry.py:13-40 on the oracle's arrays: in column
        (nx // 3, ny // 2), cells k0 = nz_solid - 1 .. nz_solid + 7, the last k > k0 with
        phi > 0.45 and |phi_k - phi_k-1| > 1e-4 sets atan(1 / slope) clamped to [10, 89] deg."""
        g = self.g
        phi = g.phi
        i, j = g.nx // 3, g.ny // 2
        k0 = max(0, self.nz_solid - 1)
        k1 = min(g.nz - 1, self.nz_solid + 8)
        best = float(self.contact_angle_deg)
        for k in range(k0, k1):
            if phi[i, j, k] > 0.45 and k > k0:
                dphi = phi[i, j, k] - phi[i, j, k - 1]
                if abs(dphi) > 1e-4:
                    slope = abs(dphi) * g.dx / max(g.dx, 1e-9)
                    best = min(89.0, max(10.0, math.degrees(math.atan(1.0 / max(slope, 1e-6)))))
        return best

    def force_diagnostics(self) -> dict:
        """Reference physics/force_diagnostics.py:36-142: max |F| of each force applied alone to
        the current state at the last arc position, then of the restored live forces."""
        g, L, m = self.g, self.lib, self.mat
        d = g.dims
        arc_i, arc_j, arc_k = self.last_arc_ijk
        saved = (g.Fx.copy(), g.Fy.copy(), g.Fz.copy())
        out = dict.fromkeys(("f_csf_max", "f_marangoni_max", "f_arc_max", "f_lorentz_max",
                             "f_buoyancy_max", "f_gravity_max", "f_gas_shear_max",
                             "f_recoil_max", "f_total_snap_max"), 0.0)

        def linf():
            return float(np.sqrt(g.Fx * g.Fx + g.Fy * g.Fy + g.Fz * g.Fz).max())

        def measure(key, fn):
            for a in (g.Fx, g.Fy, g.Fz):
                a.fill(0)
            fn()
            out[key] = linf()

        if self.enable_csf_tension:
            ct, st = cos_sin_f32(self.theta_rad)
            measure("f_csf_max", lambda: L.or_compute_csf_tension(
                ptr(g.phi), ptr(g.flags), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), self.gamma_lu,
                1 if self.enable_wetting else 0, ct, st, *d))
        if self.use_tables:
            measure("f_marangoni_max", lambda: L.or_compute_marangoni_force_variable(
                ptr(g.T), ptr(g.phi), ptr(g.f_l), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.flags),
                ptr(g.dgamma_lu_field), *d))
        else:
            measure("f_marangoni_max", lambda: L.or_compute_marangoni_force(
                ptr(g.T), ptr(g.phi), ptr(g.f_l), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.flags),
                self.dgamma_dT_lu, *d))
        if self.enable_gas_shear:
            tau = min(0.5 * 1.6 * self.gas_jet_velocity_m_s ** 2 * self.gas_shear_coeff, 5000.0)
            measure("f_gas_shear_max", lambda: L.or_apply_gas_shear_stress(
                ptr(g.Fx), ptr(g.Fy), ptr(g.phi), ptr(g.flags), arc_i, arc_j, self.sigma_cells,
                tau, g.dt, g.dx, m.rho, *d))
        measure("f_arc_max", lambda: L.or_apply_arc_pressure(
            ptr(g.Fz), ptr(g.flags), ptr(g.phi), arc_i, arc_j, arc_k,
            self.arc_pressure_sigma_m() / g.dx, self.arc_pressure_peak_pa(), g.dt, g.dx, m.rho,
            *d))
        if self.enable_recoil and self.use_recoil_cc:
            measure("f_recoil_max", lambda: L.or_apply_vapor_recoil_cc(
                ptr(g.Fz), ptr(g.T), ptr(g.phi), ptr(g.flags), arc_i, arc_j, self.sigma_cells,
                self.P_vapor_ref_Pa, self.T_boiling_K, self.L_vapor_J_kg, self.R_spec_vapor_J_kgK,
                self.recoil_accommodation, g.dt, g.dx, m.rho, *d))
        elif self.enable_recoil:
            measure("f_recoil_max", lambda: L.or_apply_vapor_recoil(
                ptr(g.Fz), ptr(g.T), ptr(g.phi), ptr(g.flags), arc_i, arc_j, self.sigma_cells,
                self.recoil_pressure_pa, m.T_liquidus, g.dt, g.dx, m.rho, *d))
        if self.enable_hydrostatic_gravity:
            measure("f_gravity_max", lambda: L.or_add_hydrostatic_gravity(
                ptr(g.Fz), ptr(g.f_l), ptr(g.flags), 1.0, self.g_lu, *d))
        measure("f_buoyancy_max", lambda: L.or_add_buoyancy(
            ptr(g.T), ptr(g.Fz), ptr(g.f_l), ptr(g.flags), self.g_lu, self.beta_T, m.T_liquidus,
            1.0, *d))
        if self.enable_lorentz:
            measure("f_lorentz_max", lambda: self.solve_lorentz(arc_i, arc_j, arc_k))
        g.Fx[...], g.Fy[...], g.Fz[...] = saved
        out["f_total_snap_max"] = linf()
        return {k: round(float(v), 6) for k, v in out.items()}

    def droplet_impact_velocity_m_s(self, m_drop):
        drop_len = 3e-3
        if self.droplet_freq > 0 and self.wire_feed_m_s > 0:
            drop_len = self.wire_feed_m_s / self.droplet_freq
        v_feed = max(self.wire_feed_m_s, 0.0)
        v_grav = math.sqrt(max(0.0, 2.0 * 9.81 * drop_len))
        dia_ref = max(self.wire_diameter_m, 1e-6)
        pinch = 0.35 * math.sqrt(max(self.welding_current_A, 0.0) / 200.0) * math.sqrt(1.2e-3 / dia_ref)
        ms = {"globular": 0.78, "spray": 1.18, "pulsed": 1.05}.get(self.transfer_mode(), 1.0)
        mass_term = 1.0
        if m_drop > 0:
            mass_term = max(0.8, min(1.25, math.sqrt(self.droplet_mass_kg() / m_drop)))
        return min(5.0, ms * max(v_feed, v_grav + pinch) * mass_term)

    def droplet_velocity_lu(self, m_drop):
        g = self.g
        v = self.droplet_impact_velocity_m_s(m_drop)
        lead = math.radians(float(self.impact_lead_angle_deg))
        dxd, dyd, _ = self._torch_dir_xyz
        dxy = math.sqrt(dxd * dxd + dyd * dyd)
        if dxy < 1e-9:
            dxd, dyd, dxy = 1.0, 0.0, 1.0
        dxd /= dxy
        dyd /= dxy
        v_xy = max(self.travel_speed_m_s, abs(v) * math.tan(lead))
        v_xy = min(v * 0.65, v_xy)
        v_z = -math.sqrt(max(v * v - v_xy * v_xy, v * v * 0.20))
        return (v_xy * dxd * g.dt / g.dx, v_xy * dyd * g.dt / g.dx, v_z * g.dt / g.dx)

    def droplet_impact_pressure_pa(self, m, r_m):
        if m <= 0:
            return 0.0
        v = self.droplet_impact_velocity_m_s(m)
        return (0.5 * m * v * v) / max(math.pi * r_m * r_m, 1e-12)

    # ---------------------------------------------------------------- step
    def _resolve_arc_k(self, arc_i, arc_j):
        g = self.g
        i0 = max(0, min(int(arc_i), g.nx - 1))
        j0 = max(0, min(int(arc_j), g.ny - 1))
        k = float(self.lib.or_surface_height_at(ptr(g.phi), ptr(g.flags), i0, j0, self.nz_solid,
                                                *g.dims))
        if k < 1.0:
            k = float(max(1, self.nz_solid - 1))
        return k

    def _ceiling(self):
        if not self.enable_enthalpy_cap:
            return
        g = self.g
        if self.use_tables:
            self.lib.or_clamp_enthalpy_ceiling_variable_cp(
                ptr(g.H), ptr(g.flags), ptr(g.cp_rho_field), self.H_liq, self.mat.T_liquidus,
                self.T_vapor_cap_K, *g.dims)
            return
        self.lib.or_clamp_enthalpy_ceiling_scalar(ptr(g.H), ptr(g.flags), self.cp_rho,
                                                  self.mat.T_solidus, self.mat.T_liquidus,
                                                  self.T_vapor_cap_K, self.L_rho, *g.dims)

    def _phase(self):
        g = self.g
        if self.use_tables:
            self.lib.or_update_phase_variable_cp(ptr(g.H), ptr(g.T), ptr(g.f_l),
                                                 ptr(g.cp_rho_field), self.L_rho,
                                                 self.mat.T_solidus, self.mat.T_liquidus,
                                                 self.H_sol, self.H_liq, *g.dims)
            return
        self.lib.or_update_phase(ptr(g.H), ptr(g.T), ptr(g.f_l), self.cp_rho, self.L_rho,
                                 self.mat.T_solidus, self.mat.T_liquidus, *g.dims)

    def _inject_arc(self, arc_i, arc_j, arc_k):
        g, L = self.g, self.lib
        pen = self.arc_penetration_m / g.dx
        en = 1 if self.arc_surface_weighting else 0
        key = self.heat_source_name.lower().replace("-", "").replace("_", "")
        if key in ("goldak", "goldak3d", "doubleellipsoid"):
            ff, fr, af, ar, b, c = self.goldak
            sgn = 1.0 if self.travel_speed_m_s >= 0.0 else -1.0
            L.or_inject_goldak_heat(ptr(g.H), ptr(g.flags), ptr(g.f_l), None, arc_i, arc_j,
                                    arc_k, self.Q_w, g.dt, g.dx ** 3, self.eta, sgn, ff, fr, af,
                                    ar, b, c, pen, en, *g.dims)
        elif key in ("conical", "conicalvolume", "cone"):
            L.or_inject_arc_heat(ptr(g.H), ptr(g.flags), ptr(g.f_l), None, arc_i, arc_j,
                                 arc_k - 1.5, self.Q_w, self.sigma_cells * 1.2, g.dt, g.dx ** 3,
                                 self.eta, pen, en, *g.dims)
        else:
            L.or_inject_arc_heat(ptr(g.H), ptr(g.flags), ptr(g.f_l), None, arc_i, arc_j, arc_k,
                                 self.Q_w, self.sigma_cells, g.dt, g.dx ** 3, self.eta, pen, en,
                                 *g.dims)

    def _droplet(self, x_m, y_m, arc_i, arc_j, arc_k):
        g, L = self.g, self.lib
        sim_time = self._step_n * g.dt
        if self.droplet_freq <= 0:
            return
        period = self.droplet_period_s()
        if sim_time - self._last_droplet_time < period:
            return
        drop_dt = sim_time - self._last_droplet_time
        self._last_droplet_time = sim_time
        drop_mass = max(0.0, self.wire_mass_flux_kg_s() * max(drop_dt, 0.0))
        drop_r = self.droplet_radius_cells(drop_mass)
        drop_vol = drop_mass / self.mat.rho
        T_drop = self.droplet_entry_temperature_K()
        L.or_inject_tracers(ptr(g.porosity_pos), ptr(g.porosity_active), ptr(g.tracer_head),
                            g.max_tracers, x_m, y_m, float(arc_k + 1) * g.dx,
                            float(self.sigma_cells * g.dx), 50, self.tracer_seed)
        foot_r = max(drop_r, self.sigma_cells * self.deposition_footprint_sigma_scale)
        r_try = foot_r
        placed = 0.0
        for _ in range(10):
            g.deposit_vol[0] = placed
            g.deposit_real[0] = placed
            L.or_feed_wire_surface(ptr(g.f_src), ptr(g.flags), ptr(g.f_l), ptr(g.phi), ptr(g.H),
                                   ptr(g.T), ptr(g.rho), arc_i, arc_j, arc_k, r_try, drop_r,
                                   drop_vol, T_drop, self.cp_rho, self.L_rho, 1.0,
                                   ptr(g.deposit_vol), ptr(g.deposit_real), g.dx ** 3, *g.dims)
            placed = float(g.deposit_real[0])
            if placed >= drop_vol * 0.98:
                break
            r_try = min(r_try + 1.5, 16.0)
        if placed < drop_vol * 0.98:
            self._deposition_overflow += 1
        self._deposited_volume_m3 += placed
        self._n_droplets_fired += 1
        if self.enable_deposition_momentum:
            vx, vy, vz = self.droplet_velocity_lu(drop_mass)
            L.or_feed_wire_momentum_impact(ptr(g.f_src), ptr(g.rho), ptr(g.ux), ptr(g.uy),
                                           ptr(g.uz), ptr(g.flags), ptr(g.f_l), arc_i, arc_j,
                                           arc_k, drop_r, vx, vy, vz, *g.dims)
            if self.enable_droplet_impact_pressure:
                L.or_apply_droplet_impact_pressure(
                    ptr(g.Fz), ptr(g.flags), ptr(g.phi), ptr(g.f_l), arc_i, arc_j, arc_k, drop_r,
                    self.droplet_impact_pressure_pa(drop_mass, drop_r * g.dx), g.dt, g.dx,
                    self.mat.rho, *g.dims)

    def step(self, torch_x_m, torch_y_m, is_welding=True, torch_z_m=None):
        # torch direction (twin.py:540-549)
        z_now = 0.0 if torch_z_m is None else float(torch_z_m)
        if self._last_torch_pos_m is not None:
            px, py, pz = self._last_torch_pos_m
            ddx, ddy, ddz = torch_x_m - px, torch_y_m - py, z_now - pz
            nrm = math.sqrt(ddx * ddx + ddy * ddy + ddz * ddz)
            if nrm > 1e-12:
                self._torch_dir_xyz = (ddx / nrm, ddy / nrm, ddz / nrm)
        self._last_torch_pos_m = (torch_x_m, torch_y_m, z_now)
        if self.enable_moving_window:
            self._maybe_shift_window(torch_x_m)
        self.coupled_step(torch_x_m - self._window_offset_x_m, torch_y_m, is_welding)

    # ---------------------------------------------------------------- Lorentz (weld_forces.py:196-281)
    def solve_lorentz(self, arc_i, arc_j, arc_k):
        """Conductivity, normalised Gaussian current source, warm-started weighted Jacobi for
        the potential (plate grounded; a cold start allows 2 n_max^2 sweeps) stopped when the
        relative L1 change of the last sweep of a 10-sweep chunk is below 1e-4, J, enclosed
        axial current per radius and slab, B_theta = mu0 I / (2 pi r), J x B on the melt."""
        g, L = self.g, self.lib
        d = g.dims
        s3 = d
        if not hasattr(g, "phi_elec"):
            for n in ("phi_elec", "phi_elec_tmp", "sigma_elec", "elec_source", "Jx", "Jy", "Jz",
                      "Bx", "By", "Bz"):
                setattr(g, n, np.zeros(s3, g.dtype))
            g.n_elec_bins = 64
            g.elec_rad_bins = np.zeros((g.nz, 64), g.dtype)
        L.or_elec_build_sigma(ptr(g.sigma_elec), ptr(g.f_l), ptr(g.flags), self.sigma_liquid_Sm,
                              self.sigma_solid_Sm, *d)
        g.elec_source.fill(0)
        L.or_elec_inject_arc_source(ptr(g.elec_source), arc_i, arc_j, arc_k, self.sigma_cells,
                                    self.welding_current_A, g.dx, *d)
        L.or_elec_normalize_source(ptr(g.elec_source), self.welding_current_A, g.dx, *d)
        cold = not self._lorentz_warm
        if cold:
            g.phi_elec.fill(0)
            self._lorentz_warm = True
        max_iters = max(int(self.lorentz_jacobi_iters), 20)
        if cold:
            n_max = max(g.nx, g.ny, g.nz)
            max_iters = max(max_iters, 2 * n_max * n_max)
        out = np.zeros(2, g.dtype)
        done, converged = 0, False
        while done < max_iters:
            for _ in range(min(10, max_iters - done)):
                L.or_elec_jacobi_step(ptr(g.phi_elec), ptr(g.phi_elec_tmp), ptr(g.sigma_elec),
                                      ptr(g.elec_source), ptr(g.flags), g.dx,
                                      self.lorentz_jacobi_omega, self.nz_solid, *d)
                L.or_elec_l1_diff(ptr(g.phi_elec_tmp), ptr(g.phi_elec), ptr(out), *d)
                g.phi_elec, g.phi_elec_tmp = g.phi_elec_tmp, g.phi_elec
                done += 1
            if float(out[0]) / max(float(out[1]), 1e-30) < self.lorentz_jacobi_tol:
                converged = True
                break
        if not converged:
            self._lorentz_unconverged += 1
        L.or_elec_compute_J(ptr(g.Jx), ptr(g.Jy), ptr(g.Jz), ptr(g.phi_elec), ptr(g.sigma_elec),
                            ptr(g.flags), g.dx, *d)
        dr = max(float(max(g.nx, g.ny)) / (2.0 * g.n_elec_bins), 0.5)
        L.or_elec_bin_axial_current(ptr(g.Jz), ptr(g.elec_rad_bins), arc_i, arc_j, g.dx, dr,
                                    g.n_elec_bins, *d)
        L.or_elec_prefix_bins(ptr(g.elec_rad_bins), g.n_elec_bins, g.nz)
        L.or_elec_B_axisymmetric(ptr(g.Bx), ptr(g.By), ptr(g.Bz), ptr(g.elec_rad_bins), arc_i,
                                 arc_j, g.dx, dr, 4.0e-7 * math.pi, g.n_elec_bins, *d)
        L.or_apply_lorentz_JxB(ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.Jx), ptr(g.Jy), ptr(g.Jz),
                               ptr(g.Bx), ptr(g.By), ptr(g.Bz), ptr(g.f_l), ptr(g.flags), g.dt,
                               g.dx, self.mat.rho, *d)

    # ---------------------------------------------------------------- moving window (twin.py:599-626)
    WINDOW_FIELDS = ("T", "H", "f_l", "phi", "T_max", "T_prev", "dT_dt", "time_above_800_s",
                     "time_above_1100_s", "time_above_solidus_s", "rho", "ux", "uy", "uz", "Fx",
                     "Fy", "Fz", "cp_rho_field", "alpha_lu_field", "dgamma_lu_field", "tau_field")

    def shift_window(self, n_shift):
        g = self.g
        arrs = [getattr(g, n) for n in self.WINDOW_FIELDS]
        table = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        self.lib.or_shift_simulation_window_x(n_shift, ptr(g.f_a), ptr(g.f_b), table,
                                              ptr(g.flags), self.T_amb, self.cp_rho, 1.0,
                                              self.nz_solid, *g.dims)

    def _maybe_shift_window(self, torch_x_world):
        g = self.g
        sim_x = torch_x_world - self._window_offset_x_m
        if sim_x < 0.62 * (g.nx - 8) * g.dx:
            return
        n_shift = max(8, g.nx // 10)
        self.shift_window(n_shift)
        self._window_offset_x_m += n_shift * g.dx
        g.porosity_pos[:, 0] -= g.dtype(n_shift * g.dx)

    def coupled_step(self, x_m, y_m, is_welding):
        g, L, m = self.g, self.lib, self.mat
        d = g.dims
        arc_i = x_m / g.dx
        arc_j = y_m / g.dx
        arc_k = self._resolve_arc_k(arc_i, arc_j)
        self.last_arc_ijk = (arc_i, arc_j, arc_k)
        g.Fx.fill(0)
        g.Fy.fill(0)
        g.Fz.fill(0)
        if self.enable_ctwd:
            self.update_ctwd()
        if is_welding:
            self._inject_arc(arc_i, arc_j, arc_k)
            self._ceiling()
            self._droplet(x_m, y_m, arc_i, arc_j, arc_k)
        if self.use_tables:
            m = self.mat
            consts = np.array([m.rho, g.dt, g.dx, m.cp, m.k, m.mu, m.dgamma_dT,
                               self.force_scale / g.dx, self.cp_rho, self.alpha_lu,
                               self.dgamma_dT_lu, g.tau, self.marangoni_scale], g.dtype)
            knots = self.table_knots.astype(g.dtype)
            L.or_refresh_thermal_properties(ptr(g.T), ptr(g.cp_rho_field), ptr(g.alpha_lu_field),
                                            ptr(g.dgamma_lu_field), ptr(g.tau_field),
                                            ptr(knots), ptr(self.table_counts),
                                            ptr(consts), 1, ptr(g.flags), *d)
            if self.mode in ("serial", "serial_thermal"):
                L.or_advect_diffuse_temperature_variable_serial(
                    ptr(g.H), ptr(g.T), ptr(g.ux), ptr(g.uy), ptr(g.uz), ptr(g.flags),
                    ptr(g.alpha_lu_field), ptr(g.cp_rho_field), 1.0, *d)
            else:
                H_in = g.H.copy()
                L.or_advect_diffuse_temperature_variable_jacobi(
                    ptr(H_in), ptr(g.H), ptr(g.T), ptr(g.ux), ptr(g.uy), ptr(g.uz),
                    ptr(g.flags), ptr(g.alpha_lu_field), ptr(g.cp_rho_field), 1.0, *d)
        elif self.mode in ("serial", "serial_thermal"):
            L.or_advect_diffuse_temperature_serial(ptr(g.H), ptr(g.T), ptr(g.ux), ptr(g.uy),
                                                   ptr(g.uz), ptr(g.flags), self.alpha_lu, 1.0,
                                                   self.cp_rho, *d)
        else:
            H_in = g.H.copy()
            L.or_advect_diffuse_temperature_jacobi(ptr(H_in), ptr(g.H), ptr(g.T), ptr(g.ux),
                                                   ptr(g.uy), ptr(g.uz), ptr(g.flags),
                                                   self.alpha_lu, 1.0, self.cp_rho, *d)
        if self.enable_heat_loss:
            L.or_apply_thermal_boundary_losses(ptr(g.H), ptr(g.T), ptr(g.flags), self.T_amb,
                                               self.h_conv, self.eps_rad,
                                               1 if self.enable_convection else 0,
                                               1 if self.enable_radiation else 0, self.cp_rho,
                                               g.dt, g.dx, self.sigma_sb, *d)
            if self.use_tables:
                L.or_clamp_enthalpy_floor(ptr(g.H), ptr(g.cp_rho_field), ptr(g.flags),
                                          self.T_amb, *d)
            else:
                L.or_clamp_enthalpy_floor_scalar(ptr(g.H), self.cp_rho, ptr(g.flags),
                                                 self.T_amb, *d)
        self._phase()
        self._ceiling()
        self._phase()
        L.or_update_T_max(ptr(g.T), ptr(g.T_max), ptr(g.flags), *d)
        L.or_update_cooling_rate(ptr(g.T), ptr(g.T_prev), ptr(g.dT_dt), ptr(g.flags), g.dt, *d)
        if self.enable_vof:
            L.or_advect_phi(ptr(g.phi_tmp), ptr(g.phi), ptr(g.ux), ptr(g.uy), ptr(g.uz),
                            ptr(g.flags), *d)
            g.phi[...] = g.phi_tmp
            L.or_reinitialize_phi(ptr(g.phi), ptr(g.flags), *d)
            if self.enable_wetting:
                ct, st = cos_sin_f32(self.theta_rad)
                if self.mode in ("serial", "serial_contact"):
                    L.or_apply_contact_angle_phi_bc_serial(ptr(g.phi), ptr(g.flags), ct, st, *d)
                else:
                    p_in = g.phi.copy()
                    L.or_apply_contact_angle_phi_bc_jacobi(ptr(p_in), ptr(g.phi), ptr(g.flags),
                                                           ct, st, *d)
            L.or_update_flags_from_phi(ptr(g.phi), ptr(g.f_l), ptr(g.flags), self.nz_solid, *d)
            L.or_sync_phi_liquid_fraction(ptr(g.phi), ptr(g.f_l), ptr(g.flags), *d)
        if self.enable_csf_tension:
            ct, st = cos_sin_f32(self.theta_rad)
            L.or_compute_csf_tension(ptr(g.phi), ptr(g.flags), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz),
                                     self.gamma_lu, 1 if self.enable_wetting else 0, ct, st, *d)
        if self.use_tables:
            L.or_compute_marangoni_force_variable(ptr(g.T), ptr(g.phi), ptr(g.f_l), ptr(g.Fx),
                                                  ptr(g.Fy), ptr(g.Fz), ptr(g.flags),
                                                  ptr(g.dgamma_lu_field), *d)
        else:
            L.or_compute_marangoni_force(ptr(g.T), ptr(g.phi), ptr(g.f_l), ptr(g.Fx),
                                         ptr(g.Fy), ptr(g.Fz), ptr(g.flags), self.dgamma_dT_lu,
                                         *d)
        if is_welding and self.enable_gas_shear:
            # shielding gas: 0.5 rho_Ar v^2 c, capped at 5 kPa
            tau = min(0.5 * 1.6 * self.gas_jet_velocity_m_s ** 2 * self.gas_shear_coeff, 5000.0)
            L.or_apply_gas_shear_stress(ptr(g.Fx), ptr(g.Fy), ptr(g.phi), ptr(g.flags), arc_i,
                                        arc_j, self.sigma_cells, tau, g.dt, g.dx, m.rho, *d)
        if is_welding:
            L.or_apply_arc_pressure(ptr(g.Fz), ptr(g.flags), ptr(g.phi), arc_i, arc_j, arc_k,
                                    self.arc_pressure_sigma_m() / g.dx,
                                    self.arc_pressure_peak_pa(), g.dt, g.dx, m.rho, *d)
            if self.enable_recoil and self.use_recoil_cc:
                L.or_apply_vapor_recoil_cc(ptr(g.Fz), ptr(g.T), ptr(g.phi), ptr(g.flags), arc_i,
                                           arc_j, self.sigma_cells, self.P_vapor_ref_Pa,
                                           self.T_boiling_K, self.L_vapor_J_kg,
                                           self.R_spec_vapor_J_kgK, self.recoil_accommodation,
                                           g.dt, g.dx, m.rho, *d)
            elif self.enable_recoil:
                L.or_apply_vapor_recoil(ptr(g.Fz), ptr(g.T), ptr(g.phi), ptr(g.flags), arc_i,
                                        arc_j, self.sigma_cells, self.recoil_pressure_pa,
                                        m.T_liquidus, g.dt, g.dx, m.rho, *d)
        if self.enable_hydrostatic_gravity:
            L.or_add_hydrostatic_gravity(ptr(g.Fz), ptr(g.f_l), ptr(g.flags), 1.0, self.g_lu, *d)
        L.or_add_buoyancy(ptr(g.T), ptr(g.Fz), ptr(g.f_l), ptr(g.flags), self.g_lu, self.beta_T,
                          m.T_liquidus, 1.0, *d)
        if is_welding and self.enable_lorentz:
            self.solve_lorentz(arc_i, arc_j, arc_k)
        if self.force_limit_lu > 0.0:
            L.or_clamp_body_force_magnitude(ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.flags),
                                            self.force_limit_lu, *d)
        if self.population_storage != "fp32":
            # the 16-bit forms hold every stored population at a representable value: values
            # written since the last step (reset, deposition, impact) are narrowed before the
            # collision reads them, and the collision's outputs as they are stored
            quantize_populations(g.f_src, self.population_storage)
            quantize_populations(g.f_dst, self.population_storage)
        if self.use_tables and self.use_variable_tau:
            L.or_collide_srt_variable_tau(ptr(g.f_src), ptr(g.f_dst), ptr(g.rho), ptr(g.ux),
                                          ptr(g.uy), ptr(g.uz), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz),
                                          ptr(g.f_l), ptr(g.flags), ptr(g.tau_field),
                                          self.C_darcy, *d)
        elif self.use_srt:
            L.or_collide_srt(ptr(g.f_src), ptr(g.f_dst), ptr(g.rho), ptr(g.ux), ptr(g.uy),
                             ptr(g.uz), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.f_l), ptr(g.flags),
                             g.tau, self.omega, 1.0, self.C_darcy, *d)
        else:
            L.or_collide_mrt(ptr(g.f_src), ptr(g.f_dst), ptr(g.rho), ptr(g.ux), ptr(g.uy),
                             ptr(g.uz), ptr(g.Fx), ptr(g.Fy), ptr(g.Fz), ptr(g.f_l), ptr(g.flags),
                             self.omega, self.omega_bulk, self.C_darcy, *d)
        if self.population_storage != "fp32":
            quantize_populations(g.f_dst, self.population_storage)
        if self.u_mach_limit_lu > 0.0:
            L.or_clamp_velocity_mach(ptr(g.ux), ptr(g.uy), ptr(g.uz), ptr(g.flags),
                                     self.u_mach_limit_lu, *d)
        L.or_stream(ptr(g.f_dst), ptr(g.f_src), ptr(g.flags), *d)
        L.or_advect_tracers(ptr(g.porosity_pos), ptr(g.porosity_active), ptr(g.ux), ptr(g.uy),
                            ptr(g.uz), ptr(g.f_l), ptr(g.flags), g.dx, g.dt, g.max_tracers, *d)
        L.or_update_time_above_T(ptr(g.T), ptr(g.flags), ptr(g.time_above_800_s),
                                 ptr(g.time_above_1100_s), ptr(g.time_above_solidus_s), g.dt,
                                 800.0 + 273.15, 1100.0 + 273.15, m.T_solidus, *d)
        g.Fx_snap[...] = g.Fx
        g.Fy_snap[...] = g.Fy
        g.Fz_snap[...] = g.Fz
        if self.enable_substrate_growth or self.enable_bead_freeze:
            if self.use_tables:
                L.or_remelt_hot_solid(ptr(g.T), ptr(g.H), ptr(g.f_l), ptr(g.phi), ptr(g.flags),
                                      ptr(g.cp_rho_field), m.T_solidus, m.T_liquidus,
                                      self.L_rho, *d)
            else:
                L.or_remelt_hot_solid_scalar(ptr(g.T), ptr(g.H), ptr(g.f_l), ptr(g.phi),
                                             ptr(g.flags), self.cp_rho, m.T_solidus,
                                             m.T_liquidus, self.L_rho, *d)
            if is_welding and self.enable_bead_freeze:
                dx_, dy_, dz_ = self._torch_dir_xyz
                lookback = max(2.0, self.trailing_solidify_lookback_mm / (g.dx * 1000.0))
                T_freeze = m.T_liquidus + self.trailing_solidify_temp_margin_K
                if self.use_tables:
                    L.or_solidify_trailing_pool(
                        ptr(g.T), ptr(g.H), ptr(g.f_l), ptr(g.phi), ptr(g.flags), ptr(g.ux),
                        ptr(g.uy), ptr(g.uz), ptr(g.cp_rho_field), arc_i, arc_j, arc_k, dx_,
                        dy_, dz_, lookback, T_freeze, m.T_solidus, *d)
                else:
                    L.or_solidify_trailing_pool_scalar(
                        ptr(g.T), ptr(g.H), ptr(g.f_l), ptr(g.phi), ptr(g.flags), ptr(g.ux),
                        ptr(g.uy), ptr(g.uz), self.cp_rho, arc_i, arc_j, arc_k, dx_, dy_, dz_,
                        lookback, T_freeze, m.T_solidus, *d)
            L.or_solidify_cooled_metal(ptr(g.T), ptr(g.f_l), ptr(g.phi), ptr(g.flags), ptr(g.ux),
                                       ptr(g.uy), ptr(g.uz), m.T_solidus,
                                       1 if (self.enable_bead_freeze or self.enable_substrate_growth) else 0,
                                       *d)
        g.swap_buffers()
        self._step_n += 1

    # ---------------------------------------------------------------- telemetry (twin.py:716-744)
    def pool_metrics(self):
        g = self.g
        out = np.zeros(8, np.float64)
        self.lib.or_telemetry_pool_reduce(ptr(g.T), ptr(g.f_l), ptr(g.phi), ptr(g.dT_dt),
                                          ptr(g.ux), ptr(g.uy), ptr(g.uz), ptr(out), *g.dims)
        n = int(out[0])
        res = {"n_liquid_cells": n, "T_global": float(out[7]), "cool": float(out[2])}
        if n > 0:
            i_c = int(round(float(np.float32(out[4])) / n))
            i_c = max(0, min(g.nx - 1, i_c))
            sect = g.f_l[i_c] > 0.5
            if sect.any():
                y_idx = np.where(sect.any(axis=1))[0]
                z_idx = np.where(sect.any(axis=0))[0]
                W_m = (y_idx[-1] - y_idx[0] + 1) * g.dx
                D_m = max(0, self.nz_solid - z_idx[0]) * g.dx
            else:
                W_m = D_m = 0.0
            res.update(T_peak=float(out[1]), i_c=i_c, pool_width_m=W_m, pool_depth_m=D_m,
                       pool_length_m=(float(out[6]) - float(out[5]) + 1.0) * g.dx,
                       u2_max=float(out[3]))
        else:
            res.update(T_peak=float(out[7]), pool_width_m=0.0, pool_depth_m=0.0,
                       pool_length_m=0.0, u2_max=0.0)
        return res


def float_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """bfloat16 words of fp32 values, rounded to nearest even (ties to the even significand);
    NaN stays a quiet NaN."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_bits_to_float(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


F16_SCALE = np.float32(32768.0)  # binary16 storage holds (f - w_q) * 2^15
F16_MAX = np.float32(65504.0)


def quantize_populations(f: np.ndarray, form: str) -> None:
    """In place: each population f[q] becomes the fp32 value its 16-bit storage form holds,
    (f - w_q) narrowed to binary16 (scaled by 2^15, saturating at +-65504) or bfloat16 (round to
    nearest even) and widened back with w_q added, every operation in fp32."""
    for q in range(Q):
        w = np.float32(W[q])
        d = (f[q] - w).astype(np.float32)
        if form == "fp16":
            s = np.clip((d * F16_SCALE).astype(np.float32), -F16_MAX, F16_MAX)
            back = (s.astype(np.float16).astype(np.float32) / F16_SCALE).astype(np.float32)
        elif form == "bf16":
            back = bf16_bits_to_float(float_to_bf16_bits(d))
        else:
            raise ValueError(form)
        f[q] = (back + w).astype(np.float32)


def cos_sin_f32(theta_rad: float) -> tuple[float, float]:
    """cos/sin of the fp32-rounded contact angle, rounded to fp32 (the kernels' ti.cos/ti.sin)."""
    t = float(np.float32(theta_rad))
    return float(np.float32(math.cos(t))), float(np.float32(math.sin(t)))
