"""The fused coupled step: one ``waam_step_fused`` call per timestep.

Same results as ``coupled_step`` (the kernel-per-reference-kernel sequence) — the parity tests
run both against the oracle — in seven launches and without host round trips on ordinary
steps (arc_k is found and used on the device). The last kernel of a step lists the cells that
end it FLUID/IFACE; the next step works through that list in the same launch as its thermal
sweep, provided no host write (``field.write_epoch``) came in between; otherwise the sweep
builds the list and one more launch consumes it. Steps on which a droplet detaches go through
the standalone arc, deposition and impact kernels first (one host read of arc_k and of the
placed volume), then the fused kernels with the arc marked as applied.

Buffer pairs: the grid holds H/H_tmp, T/T_tmp, f_l/fl_tmp, flags/flags_tmp and
sflags/sflags_tmp; after each fused step every pair swaps, as do the population buffers. The
population stream of the last step is deferred (``pending``); ``flush`` materialises it.

Population storage (``twin.population_storage``): "fp32" steps the grid's fp32 population
buffers. "fp16" / "bf16" step a pair of 16-bit buffers holding f - w_q (fp32 arithmetic in the
collision, round-to-nearest-even stores); the fp32 buffers stay the state every other kernel,
the host and the exports see. ``flush`` (``twin.sync()``) unpacks the 16-bit pair into them;
the next fused step packs them again, which returns the same 16-bit words (unpack-then-pack is
the identity on every word a pack or a collision store produces: those are fp32 differences
f - w_q, multiples of w_q's fp32 spacing, which fp32 holds exactly again once w_q is added), so
syncing does not perturb a reduced-precision run.
"""

from __future__ import annotations

import collections
import ctypes
import sys

import numpy as np
import torch

from .. import kernels
from .._lib import lib
from ..field import DeviceField, P, tensor_of, write_epoch
from ..physics import weld_forces
from ..physics.electrical_stickout import update_ctwd
from ..gpu_tables import MaterialTablesC, PropertyRefsC
from .coupled_step import (_clamp_enthalpy_ceiling, _droplet, droplet_due, droplet_plan,
                           resolve_arc_k)

_i32, _f32, _vp = ctypes.c_int32, ctypes.c_float, ctypes.c_void_p


class StepFields(ctypes.Structure):
    _fields_ = [(n, _vp) for n in (
        "f_src", "f_dst", "rho", "ux", "uy", "uz", "H_in", "H_out", "T_in", "T_out", "fl_in",
        "fl_out", "phi", "phi_tmp", "flags_in", "flags_out", "sflags_prev", "sflags_next", "Fx",
        "Fy", "Fz", "T_max", "T_prev", "dT_dt", "t800", "t1100", "tsol", "tracer_pos",
        "tracer_active", "list_thermal", "list_collide", "quiet", "cp_rho_prev", "cp_rho",
        "alpha_lu", "dgamma_lu", "tau", "scratch")] + [
            (n, _i32) for n in ("max_tracers", "nx", "ny", "nz", "layout", "pop_dtype")]

# population storage forms (WAAM_POP_*): fp32, or f - w_q in binary16 / bfloat16
POP_FORMS = {"fp32": 0, "fp16": 1, "bf16": 2}
_POP_TORCH = {"fp16": torch.float16, "bf16": torch.bfloat16}


_INT_PARAMS = ("is_welding", "collision", "arc_kind", "arc_applied", "pending_stream",
               "forces_from_fields", "store_forces", "enable_vof", "enable_wetting", "enable_csf",
               "enable_heat_loss", "enable_conv", "enable_rad", "enable_cap", "enable_hydro",
               "enable_freeze", "enable_trailing", "zero_velocity", "surface_weight", "nz_solid",
               "col_i", "col_j", "use_arc_k_override", "use_k_robot", "force_reset_all",
               "thermal_list_ready", "enable_recoil", "recoil_clausius", "enable_gas_shear",
               "quiet_skip")
_FLOAT_PARAMS = ("arc_i", "arc_j", "arc_k_override", "k_robot", "Q_w", "eta", "dt", "dx", "dx3",
                 "sigma", "pen", "g_sgn", "g_ff", "g_fr", "g_af", "g_ar", "g_b", "g_c",
                 "alpha_lu", "cp_rho", "L_rho", "T_s", "T_l", "T_cap", "T_amb", "h_conv",
                 "eps_rad", "sigma_sb", "T800", "T1100", "gamma_lu", "cos_t", "sin_t", "dgamma",
                 "p_sigma", "p_pa", "rho_phys", "g_lu", "beta", "F_cap", "u_cap", "omega_s",
                 "omega_b", "C_darcy", "dir_x", "dir_y", "dir_z", "lookback", "T_freeze",
                 "recoil_pa", "recoil_T_ref", "P_vap_ref", "T_boil", "L_vap", "R_vap", "C_acc",
                 "gas_tau")


class StepParams(ctypes.Structure):
    _fields_ = ([(n, _i32) for n in _INT_PARAMS] + [(n, _f32) for n in _FLOAT_PARAMS] +
                [("use_tables", _i32), ("variable_tau", _i32), ("H_sol_tab", _f32),
                 ("H_liq_tab", _f32), ("refs", PropertyRefsC), ("tables", MaterialTablesC)])


_f64, _u32 = ctypes.c_double, ctypes.c_uint32


class DropletC(ctypes.Structure):
    """include/waam_hip.h waam_droplet: one droplet step's deposition, tracers and impact."""

    _fields_ = [("f32_src", _vp), ("f32_dst", _vp), ("drop_vol", _f64), ("footprint_r", _f64),
                ("arc_i", _f32), ("arc_j", _f32), ("drop_r", _f32), ("T_drop", _f32),
                ("cp_rho", _f32), ("L_rho", _f32), ("rho0", _f32), ("cell_vol", _f32),
                ("vol_acc", _vp), ("real_acc", _vp), ("result", _vp), ("ledger", _vp),
                ("tracer_pos", _vp), ("tracer_active", _vp), ("tracer_head", _vp),
                ("max_tracers", _i32), ("spawn", _i32), ("seed", _u32), ("tracer_x", _f32),
                ("tracer_y", _f32), ("sigma_m", _f32), ("dx", _f64), ("momentum", _i32),
                ("pressure", _i32), ("vx", _f32), ("vy", _f32), ("vz", _f32),
                ("impact_pa", _f32), ("dt", _f32), ("dx_f", _f32), ("rho_ref", _f32)]


class _DropletArcK:
    """The arc_k a droplet step resolved on the device, read back when a caller needs it."""

    def __init__(self, g):
        self._scratch = g.scratch
        self._value = None

    def __float__(self) -> float:
        if self._value is None:
            out = ctypes.c_float()
            torch.cuda.current_stream().synchronize()
            lib().call("waam_step_droplet_arc_k", P(self._scratch), ctypes.addressof(out))
            self._value = float(out.value)
        return self._value


def _swap(g, a: str, b: str) -> None:
    fa, fb = getattr(g, a), getattr(g, b)
    setattr(g, a, fb)
    setattr(g, b, fa)


class FusedStepper:
    # steps per captured graph (WAAMTwin.step_many); graphs kept per stepper
    graph_steps = 32
    graph_cache = 4

    def __init__(self, twin):
        self.twin = twin
        g = twin.grid
        g.ensure_vof_buffers()
        for name in ("T_tmp", "fl_tmp"):
            if not hasattr(g, name):
                setattr(g, name, g.field(torch.float32))
        if not hasattr(g, "flags_tmp"):
            g.flags_tmp = g.field(torch.uint8, np_dtype=np.int32)
        if not hasattr(g, "sflags"):
            g.sflags = g.field(torch.uint8, np_dtype=np.int32)
            g.sflags_tmp = g.field(torch.uint8, np_dtype=np.int32)
        if not hasattr(g, "list_thermal"):
            # compacted FLUID/IFACE cell indices written and consumed inside each fused step
            g.list_thermal = DeviceField.zeros((g.n_cells,), torch.int32, g.device)
            g.list_collide = DeviceField.zeros((g.n_cells,), torch.int32, g.device)
        if twin.use_material_tables and not hasattr(g, "cp_rho_field_tmp"):
            # the fused step reads the previous step's cp*rho (its first ceiling) while it
            # writes its own: the two alternate, g.cp_rho_field holding the latest
            g.cp_rho_field_tmp = g.field(torch.float32)
        if not hasattr(g, "quiet"):
            # which row segments ended the last two steps quiet (all GAS, phi 0), kept by the
            # fused step itself
            nq = lib().quiet_bytes(g.nx, g.ny, g.nz, g.layout)
            g.quiet = DeviceField.zeros((nq,), torch.uint8, g.device)
        # skip the sweeps on quiet row segments (same results; the library honours it on
        # steady steps only)
        self.quiet_skip = True
        self.pending = False
        # the last fused step took its forces from the fields (an eager-arc step): the next one
        # clears every cell's forces, not only those of the previously collided cells
        self._forces_from_fields = False
        # host write count right after the last fused step: while unchanged, the thermal list
        # that step's last kernel built is this step's (nothing else has touched the flags)
        self._list_epoch = None
        # grids up to this many cells run the surface list in the thermal sweep's launch
        # (measured on MI355X: 2.1 M cells gain 3-5 us per step; at 56.6 M cells the two
        # launches take 773 us and the shared one 940 us)
        self.shared_thermal_max_cells = 1 << 23
        self._arc_out = (ctypes.c_float * 3)()
        form = getattr(twin, "population_storage", "fp32")
        if form not in POP_FORMS:
            raise ValueError(f"population_storage {form!r} (fp32 | fp16 | bf16)")
        self.pop_form, self.pop_code = form, POP_FORMS[form]
        # the C entry of a direct step: "waam_step_fused", or "waam_step_fused_noalloc" (the
        # library's own device work array instead of the grid's workspace and cell lists)
        self.entry = "waam_step_fused"
        # the 16-bit pair holds the populations (the fp32 buffers are stale) between a fused
        # step and the next flush
        self.h_live = False
        if self.pop_code:
            shape = tensor_of(g.f_a).shape
            dt = _POP_TORCH[form]
            g.hf_a = torch.zeros(shape, dtype=dt, device=g.device)
            g.hf_b = torch.zeros(shape, dtype=dt, device=g.device)
        # the swapped pairs alternate between two pointer sets: keep one struct per set
        self._fields_cache: dict[tuple, StepFields] = {}
        # captured runs of steps (waam_graph_capture), least recently used first
        self._graphs: collections.OrderedDict = collections.OrderedDict()
        self.captured_steps = 0  # steps run through captured graphs

    # ------------------------------------------------------------------ deferred stream
    def _pop_buffers(self):
        """(src, dst) population buffers the fused kernels use."""
        g = self.twin.grid
        if not self.pop_code:
            return g.f_src, g.f_dst
        return (g.hf_a, g.hf_b) if g.f_src is g.f_a else (g.hf_b, g.hf_a)

    def flush(self) -> None:
        """Materialise the deferred population stream of the last fused step and, with 16-bit
        storage, unpack the populations into the grid's fp32 buffers."""
        g = self.twin.grid
        st = torch.cuda.current_stream().cuda_stream
        if self.pending:
            src, dst = self._pop_buffers()
            lib().call("waam_stream_pending", P(src), P(dst), P(g.sflags), g.nx, g.ny, g.nz,
                       g.layout, self.pop_code, st)
            self.pending = False
        if self.h_live:
            for h, f in ((g.hf_a, g.f_a), (g.hf_b, g.f_b)):
                lib().call("waam_unpack_populations", P(h), P(f), g.n_cells, self.pop_code, st)
            self.h_live = False

    def _pack(self) -> None:
        g = self.twin.grid
        st = torch.cuda.current_stream().cuda_stream
        for f, h in ((g.f_a, g.hf_a), (g.f_b, g.hf_b)):
            lib().call("waam_pack_populations", P(f), P(h), g.n_cells, self.pop_code, st)
        self.h_live = True

    def reset(self) -> None:
        self.pending = False
        self.h_live = False
        self._list_epoch = None

    def last_arc_k(self) -> float:
        torch.cuda.current_stream().synchronize()
        lib().call("waam_step_arc_readback", P(self.twin.grid.scratch),
                   ctypes.addressof(self._arc_out))
        return float(self._arc_out[2])

    # ------------------------------------------------------------------ step
    def _fields(self) -> StepFields:
        g = self.twin.grid
        key = (id(g.H), id(g.T), id(g.f_l), id(g.flags), id(g.sflags), id(g.f_src), id(g.phi),
               id(getattr(g, "cp_rho_field", None)), self.twin.use_material_tables)
        sf = self._fields_cache.get(key)
        if sf is None:
            sf = self._fields_cache[key] = self._build_fields()
        return sf

    def _build_fields(self) -> StepFields:
        g = self.twin.grid
        sf = StepFields()
        src, dst = self._pop_buffers()
        for name, fld in (("f_src", src), ("f_dst", dst), ("rho", g.rho), ("ux", g.ux),
                          ("uy", g.uy), ("uz", g.uz), ("H_in", g.H), ("H_out", g.H_tmp),
                          ("T_in", g.T), ("T_out", g.T_tmp), ("fl_in", g.f_l),
                          ("fl_out", g.fl_tmp), ("phi", g.phi), ("phi_tmp", g.phi_tmp),
                          ("flags_in", g.flags), ("flags_out", g.flags_tmp),
                          ("sflags_prev", g.sflags), ("sflags_next", g.sflags_tmp),
                          ("Fx", g.Fx), ("Fy", g.Fy), ("Fz", g.Fz), ("T_max", g.T_max),
                          ("T_prev", g.T_prev), ("dT_dt", g.dT_dt), ("t800", g.time_above_800_s),
                          ("t1100", g.time_above_1100_s), ("tsol", g.time_above_solidus_s),
                          ("tracer_pos", g.porosity_pos), ("tracer_active", g.porosity_active),
                          ("list_thermal", g.list_thermal), ("list_collide", g.list_collide),
                          ("quiet", g.quiet), ("scratch", g.scratch)):
            setattr(sf, name, P(fld))
        if self.twin.use_material_tables:
            for name, fld in (("cp_rho_prev", g.cp_rho_field), ("cp_rho", g.cp_rho_field_tmp),
                              ("alpha_lu", g.alpha_lu_field), ("dgamma_lu", g.dgamma_lu_field),
                              ("tau", g.tau_field)):
                setattr(sf, name, P(fld))
        sf.max_tracers, sf.nx, sf.ny, sf.nz = g.max_tracers, g.nx, g.ny, g.nz
        sf.layout = g.layout
        sf.pop_dtype = self.pop_code
        return sf

    def _params(self, arc_i, arc_j, is_welding, torch_z_m) -> StepParams:
        tw, g, m = self.twin, self.twin.grid, self.twin.mat
        p = StepParams()
        src = tw.arc_source
        name = type(src).__name__
        p.is_welding = 1 if is_welding else 0
        p.collision = 1 if tw.use_srt else 0
        if tw.use_material_tables:
            # per-cell properties from the material tables (reference coupled_step.py:168-182);
            # the constants go to SOLID and GAS cells
            p.use_tables = 1
            p.H_sol_tab, p.H_liq_tab = tw.H_sol, tw.H_liq
            r = p.refs
            r.rho, r.dt, r.dx, r.force_scale = m.rho, g.dt, g.dx, tw.force_scale / g.dx
            r.cp_rho_ref, r.alpha_lu_ref = tw.cp_rho, tw.alpha_lu
            r.dgamma_lu_ref, r.tau_ref, r.marangoni_scale = tw.dgamma_dT_lu, g.tau, tw.marangoni_scale
            p.tables = tw.gpu_tables.struct
            if tw.use_variable_tau:
                p.collision, p.variable_tau = 2, 1
        p.arc_kind = 1 if name == "Goldak3D" else 0
        p.store_forces = 1 if tw.snapshot_forces else 0
        p.enable_vof = 1 if tw.enable_vof else 0
        p.enable_wetting = 1 if tw.enable_wetting else 0
        p.enable_csf = 1 if tw.enable_csf_tension else 0
        p.enable_heat_loss = 1 if tw.enable_heat_loss else 0
        p.enable_conv = 1 if tw.enable_convection else 0
        p.enable_rad = 1 if tw.enable_radiation else 0
        p.enable_cap = 1 if tw.enable_enthalpy_cap else 0
        p.enable_hydro = 1 if tw.enable_hydrostatic_gravity else 0
        p.enable_freeze = 1 if (tw.enable_substrate_growth or tw.enable_bead_freeze) else 0
        p.enable_trailing = 1 if (is_welding and tw.enable_bead_freeze) else 0
        p.zero_velocity = 1 if (tw.enable_bead_freeze or tw.enable_substrate_growth) else 0
        p.surface_weight = 1 if tw.arc_surface_weighting else 0
        p.nz_solid = tw.nz_solid
        p.col_i = max(0, min(int(arc_i), g.nx - 1))
        p.col_j = max(0, min(int(arc_j), g.ny - 1))
        k_robot = self._k_robot(torch_z_m)
        if k_robot is not None:
            p.k_robot = k_robot
            p.use_k_robot = 1
        p.arc_i, p.arc_j = arc_i, arc_j
        p.Q_w, p.eta, p.dt, p.dx, p.dx3 = tw.Q_w, tw.eta, g.dt, g.dx, g.dx ** 3
        p.pen = tw.arc_penetration_m / g.dx
        p.sigma = tw.sigma_cells
        if name == "Goldak3D":
            a_f, a_r, b, c = src.resolved_axes(tw)
            p.g_sgn = 1.0 if tw.travel_speed_m_s >= 0.0 else -1.0
            p.g_ff, p.g_fr, p.g_af, p.g_ar, p.g_b, p.g_c = src.ff, src.fr, a_f, a_r, b, c
        p.alpha_lu, p.cp_rho, p.L_rho = tw.alpha_lu, tw.cp_rho, tw.L_rho
        p.T_s, p.T_l, p.T_cap, p.T_amb = m.T_solidus, m.T_liquidus, tw.T_vapor_cap_K, tw.T_amb
        p.h_conv, p.eps_rad, p.sigma_sb = tw.h_conv, tw.eps_rad, tw.sigma_sb
        p.T800, p.T1100 = 800.0 + 273.15, 1100.0 + 273.15
        p.gamma_lu = tw.gamma_lu
        p.cos_t, p.sin_t = kernels.contact_angle_cos_sin(tw.theta_rad)
        p.dgamma = tw.dgamma_dT_lu
        p.p_sigma = weld_forces.arc_pressure_sigma_cells(tw)
        p.p_pa = weld_forces.arc_pressure_peak_pa(tw)
        p.rho_phys, p.g_lu, p.beta = m.rho, tw.g_lu, tw.beta_T
        p.F_cap = float(getattr(tw, "force_limit_lu", 0.05))
        p.u_cap = float(getattr(tw, "u_mach_limit_lu", 0.08))
        p.omega_s = tw.omega
        p.omega_b = tw.omega_bulk
        p.C_darcy = tw.C_darcy
        dir_x, dir_y, dir_z = getattr(tw, "_torch_dir_xyz", (1.0, 0.0, 0.0))
        p.dir_x, p.dir_y, p.dir_z = dir_x, dir_y, dir_z
        p.lookback = max(2.0, tw.trailing_solidify_lookback_mm / (g.dx * 1000.0))
        p.T_freeze = m.T_liquidus + tw.trailing_solidify_temp_margin_K
        p.enable_recoil = 1 if tw.enable_recoil else 0
        p.recoil_clausius = 1 if tw.use_recoil_clausius_clapeyron else 0
        p.enable_gas_shear = 1 if tw.enable_gas_shear else 0
        p.recoil_pa, p.recoil_T_ref = tw.recoil_pressure_pa, m.T_liquidus
        p.P_vap_ref, p.T_boil = tw.P_vapor_ref_Pa, tw.T_boiling_K
        p.L_vap, p.R_vap = tw.L_vapor_J_kg, tw.R_spec_vapor_J_kgK
        p.C_acc = float(getattr(tw, "recoil_accommodation", 0.54))
        p.gas_tau = weld_forces.gas_shear_tau_pa(tw)
        p.quiet_skip = 1 if self.quiet_skip else 0
        return p

    def _k_robot(self, torch_z_m) -> float | None:
        """Floor for arc_k from the robot's torch height (None: not used)."""
        tw, g = self.twin, self.twin.grid
        if not (getattr(tw, "use_torch_z", False) and torch_z_m is not None):
            return None
        z_bead = max(tw.substrate_z_m, torch_z_m - tw.ctwd_m)
        k_robot = tw.nz_solid + (z_bead - tw.substrate_z_m) / g.dx - 1.0
        return max(float(tw.nz_solid) + 0.5, min(k_robot, float(g.nz) - 2.0))

    def step(self, torch_x_m: float, torch_y_m: float, is_welding: bool,
             torch_z_m: float | None = None) -> None:
        tw, g = self.twin, self.twin.grid
        if tw.enable_lorentz:
            raise NotImplementedError("the Lorentz force solve is not built")
        arc_i = torch_x_m / g.dx
        arc_j = torch_y_m / g.dx
        p = self._params(arc_i, arc_j, is_welding, torch_z_m)
        if tw.enable_ctwd:
            # after arc_k's inputs, before this step changes the flags (reference order)
            update_ctwd(tw, g)
        conical = type(tw.arc_source).__name__ == "ConicalVolume"
        drop = None
        if is_welding and not conical and droplet_due(tw, g):
            # a droplet step on the device (waam_step_fused_droplet): arc_k, the retry loop and
            # the placed volume stay there
            p.pending_stream = 1 if self.pending else 0
            p.forces_from_fields = 1  # the impact pressure's (set again by the library)
            drop = self.droplet_struct(torch_x_m, torch_y_m, arc_i, arc_j)
            tw._last_arc_ijk = (float(arc_i), float(arc_j), _DropletArcK(g))
        # Eager arc: the conical source (whose centre sits 1.5 cells below the resolved arc_k),
        # on every welding step and with its droplets
        elif is_welding and conical:
            self.flush()
            arc_k = resolve_arc_k(tw, g, arc_i, arc_j, torch_z_m)
            kernels.clear_forces(g.Fx, g.Fy, g.Fz)
            tw.arc_source.inject(tw, g, arc_i, arc_j, arc_k)
            _clamp_enthalpy_ceiling(tw, g)
            _droplet(tw, g, torch_x_m, torch_y_m, arc_i, arc_j, arc_k)
            p.arc_applied = 1
            p.use_arc_k_override = 1
            p.arc_k_override = arc_k
            p.forces_from_fields = 1
            p.pending_stream = 0
            tw._last_arc_ijk = (float(arc_i), float(arc_j), float(arc_k))
        else:
            p.pending_stream = 1 if self.pending else 0
            # without a pending stream the previous step may not have been a fused one (reset,
            # sync, the other engine): clear every skipped cell's forces
            p.force_reset_all = 1 if (not self.pending or self._forces_from_fields) else 0
            tw._last_arc_ijk = (float(arc_i), float(arc_j), None)
        self._forces_from_fields = bool(p.forces_from_fields)
        if self._list_epoch != write_epoch():
            p.thermal_list_ready = 0
        else:
            p.thermal_list_ready = 1 if g.n_cells <= self.shared_thermal_max_cells else 2
        if self.pop_code and not self.h_live:
            self._pack()
        fields = self._fields()
        if drop is not None:
            lib().call("waam_step_fused_droplet", ctypes.addressof(p), ctypes.addressof(fields),
                       ctypes.addressof(drop), torch.cuda.current_stream().cuda_stream)
        else:
            lib().call(self.entry, ctypes.addressof(p), ctypes.addressof(fields),
                       torch.cuda.current_stream().cuda_stream)
        for a, b in (("H", "H_tmp"), ("T", "T_tmp"), ("f_l", "fl_tmp"), ("flags", "flags_tmp"),
                     ("sflags", "sflags_tmp")):
            _swap(g, a, b)
        if tw.use_material_tables:
            _swap(g, "cp_rho_field", "cp_rho_field_tmp")
        g.swap_buffers()
        self.pending = True
        self._list_epoch = write_epoch()
        tw._step_n += 1

    def droplet_struct(self, x_m: float, y_m: float, arc_i: float, arc_j: float,
                       odd: bool = False) -> DropletC:
        """The waam_droplet of a droplet detaching now (advances the twin's droplet clock and
        count, as the reference's droplet block does): the deposition, tracer and impact
        parameters the host computes in binary64 (reference coupled_step.py:110-166). odd: for
        a step an odd number of steps ahead (its population buffers swapped)."""
        tw, g = self.twin, self.twin.grid
        drop_mass, drop_r, drop_vol, T_drop, foot_r = droplet_plan(tw, g)
        d = DropletC()
        d.f32_src, d.f32_dst = P(g.f_src), P(g.f_dst)
        if odd:
            d.f32_src, d.f32_dst = d.f32_dst, d.f32_src
        d.drop_vol, d.footprint_r = float(drop_vol), float(foot_r)
        d.arc_i, d.arc_j, d.drop_r, d.T_drop = arc_i, arc_j, drop_r, T_drop
        d.cp_rho, d.L_rho, d.rho0, d.cell_vol = tw.cp_rho, tw.L_rho, 1.0, g.dx ** 3
        d.vol_acc, d.real_acc = P(g.deposit_vol_buf), P(g.deposit_real_buf)
        d.result, d.ledger = P(g.deposit_result), P(g.deposit_ledger)
        if g.max_tracers > 0:
            d.tracer_pos, d.tracer_active = P(g.porosity_pos), P(g.porosity_active)
            d.tracer_head, d.max_tracers, d.spawn = P(g.tracer_head), g.max_tracers, 50
            d.seed = int(tw.tracer_seed) & 0xFFFFFFFF
            d.tracer_x, d.tracer_y = x_m, y_m
            d.sigma_m, d.dx = float(tw.sigma_cells * g.dx), g.dx
        # counted before the impact (reference coupled_step.py:162-163): the impact velocity's
        # mass term reads the droplet period, whose jitter follows the count
        tw._n_droplets_fired += 1
        if tw.enable_deposition_momentum:
            d.momentum = 1
            d.vx, d.vy, d.vz = weld_forces.droplet_velocity_lu(tw, drop_mass)
            if tw.enable_droplet_impact_pressure:
                d.pressure = 1
                d.impact_pa = weld_forces.droplet_impact_pressure_pa(tw, drop_mass, drop_r * g.dx)
                d.dt, d.dx_f, d.rho_ref = g.dt, g.dx, tw.mat.rho
        return d

    # ------------------------------------------------------------------ captured steps
    def steady(self) -> bool:
        """Whether the next step may be a captured one: the last step was a fused step that
        left its population stream pending and its thermal list built, and nothing was written
        since (after a droplet step the captured run starts with a step that clears every
        cell's forces)."""
        return self.pending and self._list_epoch == write_epoch()

    def _odd_fields(self, fe: StepFields) -> StepFields:
        fo = StepFields.from_buffer_copy(fe)
        for a, b in (("f_src", "f_dst"), ("H_in", "H_out"), ("T_in", "T_out"),
                     ("fl_in", "fl_out"), ("flags_in", "flags_out"),
                     ("sflags_prev", "sflags_next"), ("cp_rho_prev", "cp_rho")):
            setattr(fo, a, getattr(fe, b))
            setattr(fo, b, getattr(fe, a))
        return fo

    def _graph(self, p: StepParams, n: int, kinds=None, drops=None):
        """The captured graph of n steps for these parameters, the current buffers and the run's
        plan (kinds: per-step 0 steady, 1 droplet with drops[k], 2 after a droplet; None: all
        steady)."""
        fe = self._fields()
        fo = self._odd_fields(fe)
        plan = None
        if kinds is not None and np.any(kinds):
            dl = (DropletC * n)()
            for k, d in enumerate(drops):
                if d is not None:
                    dl[k] = d
            kn = np.ascontiguousarray(kinds, dtype=np.int32)
            plan = (kn, dl)
        key = (bytes(p), bytes(fe), n) + ((plan[0].tobytes(), bytes(plan[1])) if plan else ())
        gr = self._graphs.get(key)
        if gr is None:
            while len(self._graphs) >= self.graph_cache:
                _, old = self._graphs.popitem(last=False)
                lib().call("waam_graph_destroy", old)
            h = ctypes.c_void_p()
            if plan:
                lib().call("waam_graph_capture_plan", ctypes.addressof(p), ctypes.addressof(fe),
                           ctypes.addressof(fo), n, plan[0].ctypes.data,
                           ctypes.addressof(plan[1]), ctypes.byref(h))
            else:
                lib().call("waam_graph_capture", ctypes.addressof(p), ctypes.addressof(fe),
                           ctypes.addressof(fo), n, ctypes.byref(h))
            gr = self._graphs[key] = h
        else:
            self._graphs.move_to_end(key)
        return gr

    def step_captured(self, arc_i: np.ndarray, arc_j: np.ndarray, is_welding: bool,
                      k_robot: np.ndarray | None, dirs: np.ndarray, kinds=None,
                      drops=None) -> None:
        """len(arc_i) fused steps as one graph launch: arc centres (cells, float64), robot arc_k
        floors (float32 or None) and torch directions (float32, (n, 3)) per step, and the run's
        plan (kinds, drops: see _graph); the same kernels and results as that many ``step``
        calls. The caller checks ``steady()``; with a plan, the droplet structs were made for
        these steps (droplet_struct), in order."""
        tw, g = self.twin, self.twin.grid
        n = len(arc_i)
        p = self._params(float(arc_i[0]), float(arc_j[0]), is_welding, None)
        if k_robot is not None:
            p.use_k_robot = 1
        p.pending_stream = 1
        p.force_reset_all = 0
        p.thermal_list_ready = 1 if g.n_cells <= self.shared_thermal_max_cells else 2
        # the moving parts come from the track: zero them in the graph's key
        p.arc_i = p.arc_j = p.k_robot = p.dir_x = p.dir_y = p.dir_z = 0.0
        p.col_i = p.col_j = 0
        if self.pop_code and not self.h_live:
            self._pack()
        gr = self._graph(p, n, kinds, drops)
        ai = np.ascontiguousarray(arc_i, dtype=np.float64)
        aj = np.ascontiguousarray(arc_j, dtype=np.float64)
        kr = None if k_robot is None else np.ascontiguousarray(k_robot, dtype=np.float32)
        dv = np.ascontiguousarray(dirs, dtype=np.float32).reshape(n, 3)
        lib().call("waam_graph_launch", gr, ai.ctypes.data, aj.ctypes.data,
                   None if kr is None else kr.ctypes.data, dv.ctypes.data,
                   torch.cuda.current_stream().cuda_stream)
        for _ in range(n):
            for a, b in (("H", "H_tmp"), ("T", "T_tmp"), ("f_l", "fl_tmp"),
                         ("flags", "flags_tmp"), ("sflags", "sflags_tmp")):
                _swap(g, a, b)
            if tw.use_material_tables:
                _swap(g, "cp_rho_field", "cp_rho_field_tmp")
            g.swap_buffers()
        self.pending = True
        self._forces_from_fields = kinds is not None and int(kinds[-1]) == 1
        self._list_epoch = write_epoch()
        tw._step_n += n
        self.captured_steps += n

    def close(self) -> None:
        """Destroy the captured graphs (their device track and pinned staging)."""
        graphs, self._graphs = getattr(self, "_graphs", {}), collections.OrderedDict()
        for h in graphs.values():
            lib().call("waam_graph_destroy", h)

    def __del__(self):
        if sys.is_finalizing():  # the runtime may be gone; the process frees everything
            return
        try:
            self.close()
        except Exception:
            pass
