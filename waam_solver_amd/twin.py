"""WAAMTwin — drop-in for the reference's API on the HIP implementation (reference twin.py).

Kept: the constructor keywords and their defaults (twin.py:40-287), ``from_preset`` (:360-410),
``from_job`` (:412-484, the process/heat-loss/physics-tier subset of job.py), ``reset``
(:486-531), ``step`` (:533-567), ``run`` (:627-630), ``run_path`` (:632-691) and
``get_telemetry`` (:706-821, the pool and process keys). The state lives on the GPU
(WAAMGrid); every kernel runs in libwaam_hip.
"""

from __future__ import annotations

import logging
import math
import os
import pathlib
from typing import Any

import numpy as np

from . import kernels
from .field import LAYOUT_ZYX
from .gpu_tables import MaterialGPUTables
from .grid import WAAMGrid
from .materials import MaterialProps, load_material
from .physics.arc import create_heat_source
from .physics.deposition_balance import (droplet_mass_kg, expected_deposited_mass_kg,
                                         infer_transfer_mode, wire_mass_flux_kg_s)
from .physics.bead_geometry import bead_reinforcement_height_mm, estimate_toe_angle_deg
from .physics.electrical_stickout import droplet_entry_temperature_K
from .physics.force_diagnostics import sample_force_diagnostics
from .physics.weld_forces import arc_pressure_peak_pa, droplet_impact_velocity_m_s
from .solvers.coupled_step import coupled_step, droplet_due
from .solvers.fused import FusedStepper, _DropletArcK

log = logging.getLogger("waam_solver_amd")

RHO_LATTICE = 1.0


class WAAMTwin:
    """GPU-resident WAAM melt-pool twin (LBM + enthalpy-porosity + VOF) on MI355X."""

    def __init__(self, material: str | MaterialProps = "ER70S-6", nx: int = 256, ny: int = 128,
                 nz: int = 64, dx: float = 2e-4, arc_power_W: float = 2800.0,
                 arc_efficiency: float = 0.8, high_sulphur: bool = False,
                 T_ambient: float = 300.0, use_srt: bool = True, C_darcy: float = 1.6e5,
                 arc_pressure_pa: float = 500.0, droplet_freq_hz: float = 50.0,
                 max_tracers: int | None = None, enable_heat_loss: bool = False,
                 h_conv: float = 25.0, eps_rad: float = 0.3, enable_convection: bool = True,
                 enable_radiation: bool = False, heat_source: str = "gaussian2d",
                 enable_vof: bool = False, enable_recoil: bool = False,
                 enable_csf_tension: bool = False, enable_deposition_momentum: bool = True,
                 enable_lorentz: bool = False, enable_gas_shear: bool = False,
                 enable_droplet_impact_pressure: bool = True,
                 use_recoil_clausius_clapeyron: bool = True,
                 enable_substrate_growth: bool = False, enable_moving_window: bool = False,
                 enable_wetting: bool = False, enable_hydrostatic_gravity: bool = False,
                 enable_bead_freeze: bool = False, enable_ctwd: bool = False,
                 contact_angle_deg: float | None = None, deposition_superheat_K: float = 500.0,
                 deposition_footprint_sigma_scale: float = 1.0, stickout_mm: float = 12.0,
                 ctwd_mm: float = 15.0, layer_height_mm: float = 1.2,
                 recoil_pressure_pa: float = 5000.0, welding_current_A: float = 140.0,
                 sigma_liquid_Sm: float = 1.0e6, sigma_solid_Sm: float = 1.0e5,
                 lorentz_jacobi_iters: int = 40, lorentz_jacobi_omega: float = 0.85,
                 gas_jet_velocity_m_s: float = 12.0, gas_shear_coeff: float = 1.0,
                 T_boiling_K: float = 3100.0, P_vapor_ref_Pa: float = 101325.0,
                 L_vapor_J_kg: float = 6.0e6, R_spec_vapor_J_kgK: float = 450.0,
                 travel_speed_m_s: float = 0.0, use_variable_tau: bool = True,
                 enable_enthalpy_cap: bool = True, T_vapor_cap_K: float = 3200.0,
                 arc_surface_weighting: bool = True, arc_penetration_mm: float = 2.0,
                 wire_diameter_mm: float = 1.2, arc_sigma_mm: float = 2.0,
                 bulk_tau: float | None = None, device=None, field_layout: int | None = None,
                 population_storage: str = "fp32"):
        if population_storage not in ("fp32", "fp16", "bf16"):
            raise ValueError(f"population_storage must be fp32, fp16 or bf16, got "
                             f"{population_storage!r}")
        if arc_power_W < 0:
            raise ValueError(f"arc_power_W must be >= 0, got {arc_power_W}")
        if not (0.0 < arc_efficiency <= 1.0):
            raise ValueError(f"arc_efficiency must be in (0, 1], got {arc_efficiency}")
        if wire_diameter_mm <= 0:
            raise ValueError(f"wire_diameter_mm must be > 0, got {wire_diameter_mm}")
        if arc_sigma_mm <= 0:
            raise ValueError(f"arc_sigma_mm must be > 0, got {arc_sigma_mm}")
        if contact_angle_deg is not None and not (0.0 < contact_angle_deg < 180.0):
            raise ValueError(f"contact_angle_deg must be in (0, 180), got {contact_angle_deg}")
        self.mat = material if isinstance(material, MaterialProps) else load_material(material, high_sulphur)
        max_tracers = 50_000 if max_tracers is None else max_tracers

        self.Q_w, self.eta, self.T_amb = arc_power_W, arc_efficiency, T_ambient
        self.use_srt, self.C_darcy, self.arc_pressure = use_srt, C_darcy, arc_pressure_pa
        self.arc_pressure_model = "constant"
        self.pressure_sigma_m: float | None = None
        self.physics_tier = "flow"
        self.strict_mode = False
        self._last_arc_ijk = None
        self.u_mach_limit_lu, self.force_limit_lu = 0.08, 0.05
        self.droplet_freq = droplet_freq_hz
        self.enable_heat_loss, self.h_conv, self.eps_rad = enable_heat_loss, h_conv, eps_rad
        self.enable_convection, self.enable_radiation = enable_convection, enable_radiation
        self.sigma_sb = 5.670374419e-8
        self.heat_source_name = heat_source
        self.enable_vof, self.enable_recoil = enable_vof, enable_recoil
        self.enable_csf_tension = enable_csf_tension
        self.enable_deposition_momentum = enable_deposition_momentum
        self.enable_lorentz, self.enable_gas_shear = enable_lorentz, enable_gas_shear
        # field storage: rows along x (default) or the reference's k-fastest order; the fields'
        # accessors present (nx, ny, nz) either way
        self.grid = WAAMGrid(nx, ny, nz, dx, self.mat, max_tracers=max_tracers,
                             allocate_vof=enable_vof, allocate_lorentz=enable_lorentz,
                             device=device,
                             layout=LAYOUT_ZYX if field_layout is None else field_layout)
        self.enable_droplet_impact_pressure = enable_droplet_impact_pressure
        self.use_recoil_clausius_clapeyron = use_recoil_clausius_clapeyron
        self.enable_substrate_growth = enable_substrate_growth
        self.enable_moving_window = enable_moving_window
        self.enable_wetting = enable_wetting
        self.enable_hydrostatic_gravity = enable_hydrostatic_gravity
        self.enable_bead_freeze, self.enable_ctwd = enable_bead_freeze, enable_ctwd
        self.contact_angle_deg = (contact_angle_deg if contact_angle_deg is not None
                                  else self.mat.contact_angle_deg)
        self.theta_rad = math.radians(self.contact_angle_deg)
        self.deposition_superheat_K = deposition_superheat_K
        self.deposition_footprint_sigma_scale = deposition_footprint_sigma_scale
        self.stickout_m = stickout_mm * 1e-3
        self.ctwd_nominal_m = self.ctwd_m = ctwd_mm * 1e-3
        self.layer_height_m = layer_height_mm * 1e-3
        self.rho_e_ohm_m, self.eta_stick = self.mat.rho_e_ohm_m, self.mat.eta_stick
        self.recoil_pressure_pa = recoil_pressure_pa
        self.welding_current_A = welding_current_A
        self.sigma_liquid_Sm, self.sigma_solid_Sm = sigma_liquid_Sm, sigma_solid_Sm
        self.lorentz_jacobi_iters, self.lorentz_jacobi_omega = lorentz_jacobi_iters, lorentz_jacobi_omega
        self.gas_jet_velocity_m_s, self.gas_shear_coeff = gas_jet_velocity_m_s, gas_shear_coeff
        self.T_boiling_K, self.P_vapor_ref_Pa = T_boiling_K, P_vapor_ref_Pa
        self.L_vapor_J_kg, self.R_spec_vapor_J_kgK = L_vapor_J_kg, R_spec_vapor_J_kgK
        self.travel_speed_m_s = travel_speed_m_s
        self.marangoni_scale = 1.0
        self.use_variable_tau = use_variable_tau
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
        self.plate_thickness_mm: float | None = None
        self.plate_size_mm: tuple[float, float] | None = None
        self.plate_origin_mm: tuple[float, float] | None = None
        self.nz_solid = max(1, nz // 5)
        self.preset_name: str | None = None
        self.use_torch_z = False
        self.substrate_z_m = 0.0
        self._window_offset_x_m = 0.0
        self._sim_origin_offset_x_m = 0.0
        self._sim_origin_offset_y_m = 0.0
        self._interpass_cooling_steps = 0
        self.snapshot_forces = True
        self.tracer_seed = 0
        # "fused": waam_step_fused (eight launches per step, the default); "reference": one
        # launch per reference kernel in the reference's order. Both give identical fields.
        self.engine = "fused"
        # fused steps keep the populations as fp32 ("fp32") or as f - w_q in 16 bits ("fp16",
        # "bf16": fp32 collision arithmetic, half the population traffic); every other path and
        # the host see fp32 populations (twin.sync() unpacks)
        self.population_storage = population_storage
        self._lorentz_unconverged = 0
        self._lorentz_unconverged_streak = 0
        self._lorentz_warm = False
        self._fused: FusedStepper | None = None
        # fused steps in runs go out as captured graphs (step_many, run_path)
        self.use_step_graph = True
        self.probe_recorder = None
        self._job_path: str | None = None
        self.enable_force_diagnostics = True
        self._force_diag: dict = {}
        self._reset_counters()

        g = self.grid
        mat = self.mat
        self.cp_rho = mat.rho * mat.cp
        self.L_rho = mat.rho * mat.L_fusion
        self.alpha_lu = mat.alpha * g.dt / (g.dx ** 2)
        self.omega = 1.0 / g.tau
        self.omega_bulk = 1.0 / bulk_tau if bulk_tau else self.omega
        self.arc_sigma_m = arc_sigma_mm * 1e-3
        self.sigma_cells = self.arc_sigma_m / g.dx
        self.g_lu = 9.81 * (g.dt ** 2) / g.dx
        self.beta_T = mat.beta_T
        scale_F = (g.dt ** 2) / (mat.rho * g.dx ** 2)
        self.force_scale = scale_F
        self.dgamma_dT_lu = mat.dgamma_dT * scale_F / g.dx
        self.gamma_lu = mat.gamma_0 * scale_F / g.dx
        self.gpu_tables = MaterialGPUTables(mat)
        self.use_material_tables = self.gpu_tables.enabled
        cp_sol = mat.cp_at(mat.T_solidus)
        self.H_sol = mat.rho * cp_sol * mat.T_solidus
        self.H_liq = self.H_sol + self.L_rho
        self.arc_source = create_heat_source(heat_source)

    # The deposition ledger lives on the device (grid.deposit_ledger), where the fused droplet
    # step adds each droplet's placed volume without a host read; these read and write it.
    @property
    def _deposited_volume_m3(self) -> float:
        return float(self.grid.deposit_ledger.to_numpy()[0])

    @_deposited_volume_m3.setter
    def _deposited_volume_m3(self, v: float) -> None:
        self._ledger_put(0, float(v))

    @property
    def _deposition_overflow(self) -> int:
        n = int(self.grid.deposit_ledger.to_numpy()[1])
        if n >= 20 and not getattr(self, "_warned_overflow", True):
            self._warned_overflow = True
            log.warning(f"[WAAMTwin] {n} droplets deposited less volume than the wire feed "
                        "supplies; mass_balance_ratio will drift")
        return n

    @_deposition_overflow.setter
    def _deposition_overflow(self, n: int) -> None:
        self._ledger_put(1, float(n))

    def _ledger_put(self, i: int, v: float) -> None:
        a = self.grid.deposit_ledger.to_numpy()
        a[i] = v
        self.grid.deposit_ledger.from_numpy(a)

    def _reset_counters(self) -> None:
        self._step_n = 0
        self._last_droplet_time = 0.0
        self._deposited_volume_m3 = 0.0
        self._n_droplets_fired = 0
        self._deposition_overflow = 0
        self._bead_height_m = 0.0
        self._last_torch_pos_m = None
        self._torch_dir_xyz = (1.0, 0.0, 0.0)
        self._warned_zero_wire_feed = False
        self._warned_overflow = False

    # ------------------------------------------------------------------ plate geometry
    def resolve_nz_solid(self, test_fluid_domain: bool = False) -> int:
        if test_fluid_domain:
            return 0
        g = self.grid
        headroom = max(6, g.nz // 4)
        max_solid = max(1, g.nz - headroom)
        thick = self.plate_thickness_mm
        if thick is not None and float(thick) > 0.0:
            n = int(round(float(thick) / (g.dx * 1000.0)))
            if n > max_solid:
                log.warning(f"[WAAMTwin] plate_thickness_mm={float(thick):.1f} needs {n} cells; "
                            f"clamping to {max_solid} so deposition has gas above the plate")
            return max(1, min(n, max_solid))
        return max(1, min(g.nz // 5, max_solid))

    def resolve_plate_ij(self) -> tuple[int, int, int, int]:
        g = self.grid
        dx_mm = g.dx * 1000.0
        if not self.plate_size_mm:
            return 0, g.nx, 0, g.ny
        ni = min(max(1, int(round(float(self.plate_size_mm[0]) / dx_mm))), g.nx)
        nj = min(max(1, int(round(float(self.plate_size_mm[1]) / dx_mm))), g.ny)
        if self.plate_origin_mm is not None:
            i0 = int(round(float(self.plate_origin_mm[0]) / dx_mm))
            j0 = int(round(float(self.plate_origin_mm[1]) / dx_mm))
        else:
            i0, j0 = (g.nx - ni) // 2, (g.ny - nj) // 2
        i0 = max(0, min(i0, g.nx - ni))
        j0 = max(0, min(j0, g.ny - nj))
        return i0, i0 + ni, j0, j0 + nj

    def apply_plate_geometry(self, *, plate_thickness_mm=None, plate_size_mm=None,
                             plate_origin_mm=None) -> None:
        if plate_thickness_mm is not None:
            self.plate_thickness_mm = float(plate_thickness_mm)
            self.nz_solid = self.resolve_nz_solid()
        if plate_size_mm is not None:
            self.plate_size_mm = (float(plate_size_mm[0]), float(plate_size_mm[1]))
        if plate_origin_mm is not None:
            self.plate_origin_mm = (float(plate_origin_mm[0]), float(plate_origin_mm[1]))

    # ------------------------------------------------------------------ constructors
    @classmethod
    def from_preset(cls, preset: str = "standard", material: str = "ER70S-6",
                    domain_mm=None, dx_mm: float | None = None,
                    plate_thickness_mm: float | None = None, **kwargs: Any) -> "WAAMTwin":
        from .platform import auto_grid, auto_tracer_count, check_vram_budget, resolve_preset

        cfg = resolve_preset(preset)
        tracers = auto_tracer_count(cfg.vram_budget_mb, cfg)
        dom = cfg.domain_mm if domain_mm is None else tuple(float(v) for v in domain_mm)
        dx_t = cfg.target_dx_mm if dx_mm is None else float(dx_mm)
        nx, ny, nz, dx = auto_grid(dom, dx_t, cfg.vram_budget_mb, tracers)
        check_vram_budget(nx, ny, nz, tracers, cfg.vram_budget_mb)
        twin = cls(material=material, nx=nx, ny=ny, nz=nz, dx=dx, use_srt=cfg.use_srt,
                   max_tracers=tracers, **kwargs)
        twin.preset_name = cfg.name
        if plate_thickness_mm is not None:
            twin.plate_thickness_mm = float(plate_thickness_mm)
            twin.nz_solid = twin.resolve_nz_solid()
        return twin

    @classmethod
    def from_job(cls, job_path, preset_override: str | None = None, **kwargs: Any) -> "WAAMTwin":
        from .job import apply_job_to_twin, load_job_config, resolve_plate_and_domain
        from .platform import resolve_preset

        job = load_job_config(job_path)
        if preset_override:
            job.setdefault("simulation", {})["preset"] = preset_override
        preset = job.get("simulation", {}).get("preset", "standard")
        process = job.get("process", {})
        arc_w = (float(process.get("voltage_V", 20)) * float(process.get("current_A", 140))
                 * float(process.get("power_factor", 1.0)) * float(process.get("duty_cycle", 1.0)))
        arc_eta = float(process.get("arc_efficiency", kwargs.pop("arc_efficiency", 0.8)))
        T_amb = float(process.get("T_ambient_K", kwargs.pop("T_ambient", 300.0)))
        heat = job.get("heat_loss", {})
        heat_kw = {}
        if heat:
            heat_kw = dict(enable_heat_loss=bool(heat.get("convection") or heat.get("radiation")),
                           enable_convection=bool(heat.get("convection", False)),
                           enable_radiation=bool(heat.get("radiation", False)),
                           h_conv=float(heat.get("h_conv", 25.0)),
                           eps_rad=float(heat.get("eps_rad", 0.3)))
        cfg = resolve_preset(preset)
        grid = resolve_plate_and_domain(job, cfg.domain_mm, cfg.target_dx_mm)
        twin = cls.from_preset(preset=preset, material=job.get("material", "ER70S-6"),
                               domain_mm=grid["domain_mm"], dx_mm=grid["dx_mm"],
                               plate_thickness_mm=grid["plate_thickness_mm"], arc_power_W=arc_w,
                               arc_efficiency=arc_eta, T_ambient=T_amb,
                               heat_source=str(job.get("heat_source",
                                                       kwargs.pop("heat_source", "gaussian2d"))),
                               **heat_kw, **kwargs)
        apply_job_to_twin(twin, job)
        twin.apply_plate_geometry(plate_thickness_mm=grid["plate_thickness_mm"],
                                  plate_size_mm=grid["plate_size_mm"],
                                  plate_origin_mm=grid["plate_origin_mm"])
        twin._job_config = job
        twin._job_path = str(job_path)
        if job.get("probes"):
            from .export.probes import ProbeRecorder

            twin.probe_recorder = ProbeRecorder.from_job_list(job["probes"], twin)
        return twin

    # ------------------------------------------------------------------ state
    def reset(self, T_ambient: float | None = None, test_fluid_domain: bool = False) -> None:
        T0 = T_ambient or self.T_amb
        g = self.grid
        if test_fluid_domain:
            nzs = -1
            self.nz_solid = 0
            i0, i1, j0, j1 = 0, g.nx, 0, g.ny
        else:
            self.nz_solid = self.resolve_nz_solid(False)
            nzs = self.nz_solid
            i0, i1, j0, j1 = self.resolve_plate_ij()
        kernels.init_grid(g.f_a, g.rho, g.ux, g.uy, g.uz, g.T, g.H, g.f_l, g.phi, g.flags, T0,
                          RHO_LATTICE, self.cp_rho, nzs, i0, i1, j0, j1, g.FLAG_FLUID,
                          g.FLAG_SOLID, g.FLAG_GAS)
        kernels.init_aux_fields(g.T_max, g.T_prev, g.dT_dt, g.time_above_800_s,
                                g.time_above_1100_s, g.time_above_solidus_s, g.Fx_snap,
                                g.Fy_snap, g.Fz_snap, g.Fx, g.Fy, g.Fz, T0, g.porosity_active,
                                g.tracer_head, g.max_tracers)
        if self._fused is not None:
            self._fused.reset()
        kernels.stream(g.f_a, g.f_b, g.flags, g.FLAG_SOLID, g.FLAG_GAS, g.nx, g.ny, g.nz)
        self._reset_counters()
        self._lorentz_warm = False
        g.deposit_vol_buf.fill(0.0)
        g.deposit_real_buf.fill(0.0)

    def step(self, torch_x_m: float, torch_y_m: float, is_welding: bool = True,
             torch_z_m: float | None = None) -> None:
        z_now = 0.0 if torch_z_m is None else float(torch_z_m)
        if self._last_torch_pos_m is not None:
            px, py, pz = self._last_torch_pos_m
            ddx, ddy, ddz = torch_x_m - px, torch_y_m - py, z_now - pz
            nrm = math.sqrt(ddx * ddx + ddy * ddy + ddz * ddz)
            if nrm > 1e-12:
                self._torch_dir_xyz = (ddx / nrm, ddy / nrm, ddz / nrm)
        self._last_torch_pos_m = (torch_x_m, torch_y_m, z_now)
        if (is_welding and self.droplet_freq > 0 and self.wire_feed_m_s <= 0.0
                and not self._warned_zero_wire_feed):
            self._warned_zero_wire_feed = True
            log.warning("[WAAMTwin] welding step with wire_feed_m_s == 0: no metal is deposited")
        if self.enable_moving_window:
            self._maybe_shift_window(torch_x_m)
        sim_x = torch_x_m - self._window_offset_x_m - self._sim_origin_offset_x_m
        sim_y = torch_y_m - self._sim_origin_offset_y_m
        if self.engine == "fused" and self.enable_lorentz:
            # the Lorentz solve (a Jacobi iteration over the whole grid that dominates such a
            # step) runs through the reference-order kernels, which give the same fields
            self.sync()
            coupled_step(self, sim_x, sim_y, is_welding, torch_z_m)
        elif self.engine == "fused":
            if self._fused is None:
                self._fused = FusedStepper(self)
            self._fused.step(sim_x, sim_y, is_welding, torch_z_m)
        elif self.engine == "reference":
            self.sync()
            coupled_step(self, sim_x, sim_y, is_welding, torch_z_m)
        else:
            raise ValueError(f"unknown engine {self.engine!r} (fused | reference)")
        if self.probe_recorder is not None:
            self.probe_recorder.record_step(self)
        self._check_strict_mode()

    def step_many(self, torch_xs, torch_ys, is_welding: bool = True, torch_zs=None) -> None:
        """``step`` for each position in turn, with the same results. With the fused engine,
        runs of ``FusedStepper.graph_steps`` steady steps (no droplet, no window shift, no
        per-step host work such as probes or CTWD) go to the GPU as one captured graph."""
        xs = [float(v) for v in torch_xs]
        ys = [float(v) for v in torch_ys]
        zs = None if torch_zs is None else list(torch_zs)
        if len(ys) != len(xs) or (zs is not None and len(zs) != len(xs)):
            raise ValueError("step_many: position lists differ in length")
        s, n = 0, len(xs)
        while s < n:
            m = self._captured_run(xs, zs, s, is_welding)
            if m:
                self._step_captured(xs[s:s + m], ys[s:s + m],
                                    None if zs is None else zs[s:s + m], is_welding)
                s += m
            else:
                self.step(xs[s], ys[s], is_welding, None if zs is None else zs[s])
                s += 1

    def _captured_run(self, xs, zs, s: int, is_welding: bool) -> int:
        """Steps from position s that may go out as one captured graph (0: step directly)."""
        st = self._fused
        if not self.use_step_graph or self.engine != "fused" or st is None:
            return 0
        m = st.graph_steps
        if m <= 0 or s + m > len(xs) or not st.steady():
            return 0
        if (self.enable_lorentz or self.enable_ctwd
                or self.probe_recorder is not None
                or type(self.arc_source).__name__ == "ConicalVolume"):
            return 0
        if (is_welding and self.droplet_freq > 0 and self.wire_feed_m_s <= 0.0
                and not self._warned_zero_wire_feed):
            return 0
        if self.use_torch_z and zs is not None and any(z is None for z in zs[s:s + m]):
            return 0
        g = self.grid
        if self.enable_moving_window:
            trigger = 0.62 * (g.nx - 8) * g.dx
            if any(x - self._window_offset_x_m >= trigger for x in xs[s:s + m]):
                return 0
        return m

    def _step_captured(self, xs, ys, zs, is_welding: bool) -> None:
        """The host side of ``step`` for each position (torch direction, last position, arc
        centre), then one captured launch of all the steps."""
        g, st = self.grid, self._fused
        m = len(xs)
        base_n = self._step_n
        # the run's plan: 0 a steady step, 1 a droplet step (its deposition parameters in
        # drops[k]), 2 the step after a droplet step
        kinds = np.zeros(m, dtype=np.int32)
        drops = [None] * m
        arc_i = np.empty(m, dtype=np.float64)
        arc_j = np.empty(m, dtype=np.float64)
        dirs = np.empty((m, 3), dtype=np.float32)
        use_kr = self.use_torch_z and zs is not None
        k_robot = np.empty(m, dtype=np.float32) if use_kr else None
        for k in range(m):
            x, y = xs[k], ys[k]
            z = None if zs is None else zs[k]
            z_now = 0.0 if z is None else float(z)
            if self._last_torch_pos_m is not None:
                px, py, pz = self._last_torch_pos_m
                ddx, ddy, ddz = x - px, y - py, z_now - pz
                nrm = math.sqrt(ddx * ddx + ddy * ddy + ddz * ddz)
                if nrm > 1e-12:
                    self._torch_dir_xyz = (ddx / nrm, ddy / nrm, ddz / nrm)
            self._last_torch_pos_m = (x, y, z_now)
            arc_i[k] = (x - self._window_offset_x_m - self._sim_origin_offset_x_m) / g.dx
            arc_j[k] = (y - self._sim_origin_offset_y_m) / g.dx
            dirs[k] = self._torch_dir_xyz
            if use_kr:
                k_robot[k] = st._k_robot(z)
            if (k == 0 and st._forces_from_fields) or (k > 0 and kinds[k - 1] == 1):
                kinds[k] = 2
            if is_welding and self.droplet_freq > 0:
                # the droplet clock as the direct steps would advance it
                self._step_n = base_n + k
                if droplet_due(self, g):
                    kinds[k] = 1
                    drops[k] = st.droplet_struct(
                        x - self._window_offset_x_m - self._sim_origin_offset_x_m,
                        y - self._sim_origin_offset_y_m, arc_i[k], arc_j[k], odd=bool(k & 1))
                self._step_n = base_n
        st.step_captured(arc_i, arc_j, is_welding, k_robot, dirs, kinds, drops)
        if kinds[-1] == 1:
            self._last_arc_ijk = (float(arc_i[-1]), float(arc_j[-1]), _DropletArcK(g))
        else:
            self._last_arc_ijk = (float(arc_i[-1]), float(arc_j[-1]), None)
        self._check_strict_mode()

    def _maybe_shift_window(self, torch_x_world: float) -> None:
        """Shift the domain by max(8, nx // 10) cells once the torch passes 62 % of the usable
        length (reference twin.py:599-626); the torch offset and tracer positions follow."""
        g = self.grid
        sim_x = torch_x_world - self._window_offset_x_m
        trigger = 0.62 * (g.nx - 8) * g.dx
        if sim_x < trigger:
            return
        n_shift = max(8, g.nx // 10)
        self.sync()  # the deferred population stream must land before the shift
        kernels.shift_simulation_window_x(
            n_shift, g.f_a, g.f_b, g.T, g.H, g.f_l, g.phi, g.flags, g.T_max, g.T_prev, g.dT_dt,
            g.time_above_800_s, g.time_above_1100_s, g.time_above_solidus_s, g.rho, g.ux, g.uy,
            g.uz, g.Fx, g.Fy, g.Fz, g.cp_rho_field, g.alpha_lu_field, g.dgamma_lu_field,
            g.tau_field, self.T_amb, self.cp_rho, RHO_LATTICE, self.nz_solid, g.FLAG_SOLID,
            g.FLAG_GAS, g.FLAG_FLUID, g.nx, g.ny, g.nz)
        self._window_offset_x_m += n_shift * g.dx
        pos = g.porosity_pos.to_numpy()
        pos[:, 0] -= n_shift * g.dx
        g.porosity_pos.from_numpy(pos)
        log.info(f"[WAAMTwin] moving window shift -{n_shift} cells, "
                 f"offset={self._window_offset_x_m * 1e3:.2f} mm")

    def sync(self) -> None:
        """Materialise deferred device work (the fused engine's last population stream).

        Fields other than the two population buffers are always current; call this before
        reading ``grid.f_src`` / ``grid.f_dst`` on the host or handing them to other kernels.
        """
        if self._fused is not None:
            self._fused.flush()

    def _check_strict_mode(self) -> None:
        """Reference twin.py:569-597: a Lorentz solve that failed for more than 50 consecutive
        steps, or a drifting deposited-mass ledger, stops the run."""
        if not self.strict_mode:
            return
        if self.enable_lorentz and self._lorentz_unconverged_streak > 50:
            raise RuntimeError(f"[strict_mode] Lorentz Jacobi failed to converge for "
                               f"{self._lorentz_unconverged_streak} consecutive steps")
        if self._n_droplets_fired < 20:
            return
        ratio = (self._deposited_volume_m3 * self.mat.rho) / max(
            self._n_droplets_fired * droplet_mass_kg(self), 1e-12)
        if not (0.95 <= ratio <= 1.05):
            raise RuntimeError(f"[strict_mode] mass_balance_ratio={ratio:.3f} outside [0.95, 1.05] "
                               f"after {self._n_droplets_fired} droplets "
                               f"(overflow_count={self._deposition_overflow})")

    def run(self, n_steps: int, torch_x_m: float = 0.0, torch_y_m: float = 0.0,
            is_welding: bool = True) -> None:
        for _ in range(n_steps):
            self.step(torch_x_m, torch_y_m, is_welding)

    def run_path(self, job, n_steps: int | None = None, is_welding: bool = True,
                 interpass_steps: int = 0) -> None:
        """Follow the job's torch path at travel_speed_m_s (reference twin.py:632-691)."""
        from .job import load_job_config, parse_torch_path
        from .torch_path import TorchPathDriver, clamp_torch_to_domain

        if not isinstance(job, dict):
            job = load_job_config(job)
        waypoints = parse_torch_path(job)
        g = self.grid
        if not waypoints:
            waypoints = [(0.01, (g.ny // 2) * g.dx, 0.0)]
        driver = TorchPathDriver(waypoints, self.travel_speed_m_s)
        if n_steps is None:
            n_steps = max(400, int(driver.total_length / max(self.travel_speed_m_s * g.dt, 1e-12)))
        interpass = int(interpass_steps) if interpass_steps else int(self._interpass_cooling_steps or 0)
        prev_seg = -1
        if driver.segments:
            s0 = driver.segments[0]
            d = (s0.x1 - s0.x0, s0.y1 - s0.y0, s0.z1 - s0.z0)
            nrm = math.sqrt(sum(v * v for v in d))
            if nrm > 1e-12:
                self._torch_dir_xyz = tuple(v / nrm for v in d)

        def _clamp(x_m, y_m):
            off = self._window_offset_x_m
            cx, cy = clamp_torch_to_domain(x_m - off, y_m, g.nx, g.ny, g.dx)
            return cx + off, cy

        # without the moving window the clamp does not depend on the steps taken, so the
        # positions are planned first and runs of equal welding state go through step_many
        plan: list[tuple[float, float, bool, float]] = []

        def _go(x_m, y_m, welding, z_m):
            if self.enable_moving_window:
                cx, cy = _clamp(x_m, y_m)
                self.step(cx, cy, welding, torch_z_m=z_m)
            else:
                plan.append((*_clamp(x_m, y_m), welding, z_m))

        for step, x, y, z in driver.positions_for_steps(n_steps, g.dt):
            seg = driver.segment_index_at_distance(driver.distance_at_step(step, g.dt))
            if interpass > 0 and seg > prev_seg and prev_seg >= 0:
                park = driver.segment_end(prev_seg)
                px, py, pz = park if park else (x, y, z)
                for _ in range(interpass):
                    _go(px, py, False, pz)
            prev_seg = seg
            _go(x, y, is_welding, z)
        s = 0
        while s < len(plan):
            e = s
            while e < len(plan) and plan[e][2] == plan[s][2]:
                e += 1
            run = plan[s:e]
            self.step_many([r[0] for r in run], [r[1] for r in run], run[0][2],
                           [r[3] for r in run])
            s = e

    # ------------------------------------------------------------------ telemetry
    def pool_reduce(self) -> np.ndarray:
        """The eight pool aggregates (reference kernels.py:3424-3468) as float64 on the host."""
        g = self.grid
        kernels.telemetry_pool_reduce_into(g.T, g.f_l, g.phi, g.dT_dt, g.ux, g.uy, g.uz,
                                           g.telem_out, g.scratch)
        return g.telem_out.to_numpy()

    def get_telemetry(self) -> dict:
        g = self.grid
        out = self.pool_reduce()
        n_liquid = int(out[0])
        if n_liquid > 0:
            T_peak, peak_cool = float(out[1]), float(out[2])
            i_c = int(round(float(np.float32(out[4])) / n_liquid))
            i_c = max(0, min(g.nx - 1, i_c))
            kernels.extract_fl_yz_slice(g.f_l, g.telem_fl_slice, i_c, g.ny, g.nz)
            sect = g.telem_fl_slice.to_numpy() > 0.5
            if sect.any():
                y_idx = np.where(sect.any(axis=1))[0]
                z_idx = np.where(sect.any(axis=0))[0]
                pool_width_m = (y_idx[-1] - y_idx[0] + 1) * g.dx
                pool_depth_m = max(0, self.nz_solid - z_idx[0]) * g.dx
            else:
                pool_width_m = pool_depth_m = 0.0
            pool_length_m = (float(out[6]) - float(out[5]) + 1.0) * g.dx
            u_max_phys = float(np.sqrt(float(out[3]))) * g.dx / g.dt
        else:
            T_peak, peak_cool = float(out[7]), float(out[2])
            pool_width_m = pool_depth_m = pool_length_m = u_max_phys = 0.0
        active = g.porosity_active.to_numpy()
        n_trapped, n_active = int((active == 2).sum()), int((active == 1).sum())
        bead_h = bead_reinforcement_height_mm(self, g)
        toe_deg = estimate_toe_angle_deg(self, g) if self.enable_wetting else 0.0
        force_diag = dict(self._force_diag or {})
        if self.enable_force_diagnostics and self._step_n > 0:
            try:
                force_diag = sample_force_diagnostics(self)
                self._force_diag = force_diag
            except Exception as exc:  # reported in the telemetry, as the reference does
                force_diag = {"error": str(exc)}
        dep_mass = self._deposited_volume_m3 * self.mat.rho
        m_drop = droplet_mass_kg(self)
        exp_drop = self._n_droplets_fired * m_drop
        return {
            "step": self._step_n,
            "sim_time_ms": round(self._step_n * g.dt * 1000, 4),
            "peak_temp_K": round(T_peak, 1),
            "peak_temp_C": round(T_peak - 273.15, 1),
            "peak_cooling_rate_Ks": round(peak_cool, 1),
            "pool_width_mm": round(pool_width_m * 1000, 3),
            "pool_depth_mm": round(pool_depth_m * 1000, 3),
            "pool_length_mm": round(pool_length_m * 1000, 3),
            "n_liquid_cells": n_liquid,
            "marangoni_vel_ms": round(u_max_phys, 4),
            "material_status": self.mat.status,
            "material_name": self.mat.name,
            "heat_source": self.heat_source_name,
            "material_tables": self.use_material_tables,
            "vof_enabled": self.enable_vof,
            "travel_speed_mm_s": round(self.travel_speed_m_s * 1000, 3),
            "variable_tau": self.use_variable_tau and self.use_material_tables,
            "porosity_pct": round(100.0 * n_trapped / max(n_trapped + n_active, 1), 3),
            "n_trapped_tracers": n_trapped,
            "window_offset_x_mm": round(self._window_offset_x_m * 1000, 3),
            "deposited_mass_g": round(dep_mass * 1000, 4),
            "expected_wire_mass_g": round(expected_deposited_mass_kg(self) * 1000, 4),
            "expected_drop_mass_g": round(exp_drop * 1000, 4),
            "n_droplets_fired": self._n_droplets_fired,
            "droplet_mass_mg": round(m_drop * 1e6, 3),
            "droplet_transfer_mode": infer_transfer_mode(self),
            "droplet_impact_velocity_ms": round(droplet_impact_velocity_m_s(self), 4),
            "mass_balance_ratio": round(dep_mass / max(exp_drop, 1e-12), 3),
            "wire_mass_flux_g_s": round(wire_mass_flux_kg_s(self) * 1000, 4),
            "T_vapor_cap_K": self.T_vapor_cap_K,
            "arc_surface_weighting": self.arc_surface_weighting,
            "enable_enthalpy_cap": self.enable_enthalpy_cap,
            "enable_wetting": self.enable_wetting,
            "enable_hydrostatic_gravity": self.enable_hydrostatic_gravity,
            "enable_bead_freeze": self.enable_bead_freeze,
            "enable_ctwd": self.enable_ctwd,
            "contact_angle_deg": round(self.contact_angle_deg, 1),
            "bead_height_mm": round(bead_h, 3),
            "toe_angle_deg": round(toe_deg, 1),
            "ctwd_mm": round(self.ctwd_m * 1000, 3),
            "T_drop_K": round(droplet_entry_temperature_K(self), 1),
            "deposition_overflow_count": self._deposition_overflow,
            "physics_tier": self.physics_tier,
            "strict_mode": bool(self.strict_mode),
            "arc_pressure_model": self.arc_pressure_model,
            "arc_pressure_peak_pa": round(arc_pressure_peak_pa(self), 1),
            "enable_lorentz": self.enable_lorentz,
            "enable_gas_shear": self.enable_gas_shear,
            "lorentz_unconverged_count": int(self._lorentz_unconverged or 0),
            "lorentz_unconverged_streak": int(self._lorentz_unconverged_streak or 0),
            "force_diagnostics": force_diag,
        }

    # ------------------------------------------------------------------ export
    def export_vtk(self, path: str = "weld_pool.vti"):
        """Core + derived tiers (reference twin.py:823-825)."""
        return self.export_vtk_full(path, tiers=(0, 3), crop_liquid=False)

    def export_vtk_full(self, path: str = "weld_pool_full.vti", tiers=(0, 1, 2, 3),
                        crop_liquid: bool = False):
        from .export import vtk_io

        return vtk_io.export_volume(self, path, tiers=tuple(vtk_io.TIER_MAP.get(t, t)
                                                            for t in tiers),
                                    crop_liquid=crop_liquid)

    def export_surface_vtk(self, path: str = "bead_surface.vtp"):
        from .export.vtk_io import export_surface

        return export_surface(self, path, include_kappa=True)

    def export_tracers_vtk(self, path: str = "tracers.vtp"):
        from .export.vtk_io import export_tracers

        return export_tracers(self, path)

    def export_research_bundle(self, out_dir, tag: str | None = None, tiers=(0, 1, 3),
                               include_surface: bool = True, include_tracers: bool = True,
                               crop_liquid: bool = False) -> dict[str, str]:
        from .export.bundle import export_research_bundle

        return export_research_bundle(self, out_dir, tag=tag, tiers=tiers,
                                      include_surface=include_surface,
                                      include_tracers=include_tracers, crop_liquid=crop_liquid,
                                      job_path=self._job_path)

    def export_research_sequence(self, *args, **kwargs):
        from .export.bundle import export_research_sequence

        return export_research_sequence(self, *args, job_path=self._job_path, **kwargs)
