"""The differentiation form: gradients of the thermal sub-step through many steps.

Forward: the coupled step's thermal chain with frozen flags (and phi) and zero velocity, on a
copy of a twin's state (reference solvers/coupled_step.py:97-160, the flow and the free surface
held): Gaussian2D arc injection, enthalpy ceiling, conduction from the step's starting
temperature, boundary losses, enthalpy floor, phase map, ceiling, phase map — the product's own
kernels, one launch each. Parameters theta = (eta, s_h, s_sigma): the arc efficiency, a factor
on both loss coefficients (the calibration overlay's heat-loss factor) and a factor on the arc
radius (its sigma scale).

Backward: hand-written adjoint kernels (``csrc/adjoint.hip``) step the adjoint state
lambda = dL/dH back through each step and accumulate dL/dtheta in binary64. The forward keeps
the enthalpy every ``checkpoint_every`` steps; the backward recomputes each segment from its
checkpoint (keeping the step's enthalpy after the arc, before and after the floor, and its
starting f_l), then runs the segment's adjoint steps in reverse. ``ThermalRolloutFn`` wraps the
pair as a ``torch.autograd.Function`` of theta returning the final (T, f_l), so any loss written
in torch differentiates through it; ``fit_calibration`` runs gradient descent on it.

Derivatives at the piecewise points are one-sided (a clamped value has derivative 0; the phase
map's slope is the one of the branch the enthalpy is on); f_l enters the arc's surface weight
only through a step at 0.55, whose derivative is 0.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import kernels
from ._lib import lib
from .field import P, tensor_of
from .physics import weld_forces
from .solvers.coupled_step import resolve_arc_k

_f = ctypes.c_float


class AdjConsts(ctypes.Structure):
    _fields_ = [(n, _f) for n in (
        "cp_rho", "L_rho", "T_s", "T_l", "T_cap", "T_floor", "alpha_lu", "T_amb", "h_conv0",
        "eps_rad0", "sigma_sb", "dt", "dx", "loss_scale", "arc_i", "arc_j", "arc_k", "sigma",
        "sigma0", "Q_w", "eta", "dx3", "pen")] + [(n, ctypes.c_int32) for n in (
            "enable_cap", "enable_loss", "enable_arc", "enable_conv", "enable_rad",
            "surface_weight")]


class ThermalRollout:
    """n_steps of the thermal sub-step from the twin's current state along ``path``
    (step -> torch (x_m, y_m)), differentiable in theta."""

    def __init__(self, twin, n_steps: int, path, checkpoint_every: int = 50,
                 is_welding: bool = True):
        if type(twin.arc_source).__name__ != "Gaussian2D":
            raise NotImplementedError("the differentiation form covers the Gaussian2D arc")
        twin.sync()
        self.tw, g = twin, twin.grid
        self.n_steps, self.K = int(n_steps), max(1, int(checkpoint_every))
        self.path, self.is_welding = path, bool(is_welding)
        f = lambda: g.field(torch.float32)  # noqa: E731
        self.H, self.T, self.fl, self.Htmp = f(), f(), f(), f()
        # the starting state is the twin's enthalpy with T and f_l from the phase map, as at
        # every later step boundary (and as the checkpoints restore it)
        self.H0, self.T0, self.fl0 = f(), f(), f()
        tensor_of(self.H0).copy_(tensor_of(g.H))
        tensor_of(self.H).copy_(tensor_of(g.H))
        self._phase()
        tensor_of(self.T0).copy_(tensor_of(self.T))
        tensor_of(self.fl0).copy_(tensor_of(self.fl))
        self.zero_u = f()
        # the rollout's own copies of the frozen flags and phi: the twin may step (and the fused
        # step updates phi in place and swaps its flag buffers) between forward and backward
        self.flags = g.field(torch.uint8, np_dtype=np.int32)
        self.phi = f()
        tensor_of(self.flags).copy_(tensor_of(g.flags))
        tensor_of(self.phi).copy_(tensor_of(g.phi))
        self.h_conv0, self.eps0 = float(twin.h_conv), float(twin.eps_rad)
        self.sigma0 = float(twin.sigma_cells)
        x0, y0 = path(0)
        self.arc_k = resolve_arc_k(twin, g, x0 / g.dx, y0 / g.dx)
        self.norm_buf = torch.zeros((), dtype=torch.float32, device=g.device)
        self.ckpt: list[tuple[torch.Tensor, torch.Tensor, torch.Tensor]] = []
        self.theta = None

    # ------------------------------------------------------------------ forward
    def _step(self, n: int, theta, save=None) -> None:
        tw, g = self.tw, self.tw.grid
        eta, s_h, s_sig = (float(v) for v in theta)
        H, T, fl = self.H, self.T, self.fl
        x, y = self.path(n)
        if self.is_welding:
            kernels.inject_arc_heat(H, self.flags, self.phi, fl, self.norm_buf, x / g.dx, y / g.dx,
                                    self.arc_k, tw.Q_w, self.sigma0 * s_sig, g.dt, g.dx ** 3, eta,
                                    tw.arc_penetration_m / g.dx,
                                    1 if tw.arc_surface_weighting else 0, g.FLAG_SOLID,
                                    g.FLAG_GAS, scratch=g.scratch)
        if save is not None:
            save["A"].copy_(tensor_of(H))
        self._ceiling(H)
        kernels.advect_diffuse_temperature_into(H, self.Htmp, T, self.zero_u, self.zero_u,
                                                self.zero_u, self.flags, tw.alpha_lu, 1.0,
                                                tw.cp_rho, g.nx, g.ny, g.nz)
        self.H, self.Htmp = self.Htmp, self.H
        H = self.H
        if tw.enable_heat_loss:
            kernels.apply_thermal_boundary_losses(
                H, T, self.flags, tw.T_amb, self.h_conv0 * s_h, self.eps0 * s_h,
                1 if tw.enable_convection else 0, 1 if tw.enable_radiation else 0, tw.cp_rho,
                g.dt, g.dx, tw.sigma_sb, g.FLAG_SOLID, g.FLAG_GAS, g.nx, g.ny, g.nz)
        if save is not None:
            save["E"].copy_(tensor_of(H))
        if tw.enable_heat_loss:
            kernels.clamp_enthalpy_floor_scalar(H, tw.cp_rho, self.flags, tw.T_amb, g.FLAG_GAS)
        if save is not None:
            save["F"].copy_(tensor_of(H))
        self._phase()
        self._ceiling(H)
        self._phase()

    def _ceiling(self, H) -> None:
        tw, g = self.tw, self.tw.grid
        if tw.enable_enthalpy_cap:
            kernels.clamp_enthalpy_ceiling_scalar(H, self.flags, tw.cp_rho, tw.mat.T_solidus,
                                                  tw.mat.T_liquidus, tw.T_vapor_cap_K, tw.L_rho,
                                                  g.FLAG_GAS)

    def _phase(self) -> None:
        tw = self.tw
        kernels.update_phase(self.H, self.T, self.fl, tw.cp_rho, tw.L_rho, tw.mat.T_solidus,
                             tw.mat.T_liquidus)

    def _load(self, H) -> None:
        tensor_of(self.H).copy_(H)
        self._phase()

    def forward(self, theta) -> tuple[torch.Tensor, torch.Tensor]:
        """Final (T, f_l) (storage order) after n_steps with parameters theta."""
        self.theta = tuple(float(v) for v in theta)
        self.ckpt = []
        tensor_of(self.H).copy_(tensor_of(self.H0))
        tensor_of(self.T).copy_(tensor_of(self.T0))
        tensor_of(self.fl).copy_(tensor_of(self.fl0))
        for n in range(self.n_steps):
            if n % self.K == 0:
                self.ckpt.append(tensor_of(self.H).clone())
            self._step(n, self.theta)
        self.H_final = tensor_of(self.H).clone()
        return tensor_of(self.T).clone(), tensor_of(self.fl).clone()

    # ------------------------------------------------------------------ backward
    def _consts(self, n: int) -> AdjConsts:
        tw, g = self.tw, self.tw.grid
        eta, s_h, s_sig = self.theta
        x, y = self.path(n)
        m = tw.mat
        c = AdjConsts(cp_rho=tw.cp_rho, L_rho=tw.L_rho, T_s=m.T_solidus, T_l=m.T_liquidus,
                      T_cap=tw.T_vapor_cap_K, T_floor=tw.T_amb, alpha_lu=tw.alpha_lu,
                      T_amb=tw.T_amb, h_conv0=self.h_conv0, eps_rad0=self.eps0,
                      sigma_sb=tw.sigma_sb, dt=g.dt, dx=g.dx, loss_scale=s_h, arc_i=x / g.dx,
                      arc_j=y / g.dx, arc_k=self.arc_k, sigma=self.sigma0 * s_sig,
                      sigma0=self.sigma0, Q_w=tw.Q_w, eta=eta, dx3=g.dx ** 3,
                      pen=tw.arc_penetration_m / g.dx)
        c.enable_cap = 1 if tw.enable_enthalpy_cap else 0
        c.enable_loss = 1 if tw.enable_heat_loss else 0
        c.enable_arc = 1 if self.is_welding else 0
        c.enable_conv = 1 if tw.enable_convection else 0
        c.enable_rad = 1 if tw.enable_radiation else 0
        c.surface_weight = 1 if tw.arc_surface_weighting else 0
        return c

    def backward(self, grad_T: torch.Tensor, grad_fl: torch.Tensor) -> np.ndarray:
        """dL/dtheta (float64) given dL/dT and dL/df_l at the final step."""
        if self.theta is None:
            raise RuntimeError("backward before forward")
        tw, g = self.tw, self.tw.grid
        st = torch.cuda.current_stream().cuda_stream
        dev = tensor_of(self.H).device
        n_cells = g.n_cells
        shp = tensor_of(self.H).shape
        lam = torch.empty(shp, dtype=torch.float32, device=dev)
        lam2 = torch.empty_like(lam)
        lamC = torch.empty_like(lam)
        grads = torch.zeros(3, dtype=torch.float64, device=dev)
        sums = torch.zeros(2, dtype=torch.float64, device=dev)
        c_last = self._consts(self.n_steps - 1)
        # P() hands over a bare address: the converted seeds are held in locals until the
        # kernels that read them are queued (a temporary would be freed, and its memory reused,
        # before the call)
        gT = grad_T.float().contiguous()
        gfl = grad_fl.float().contiguous()
        lib().call("waam_adj_terminal", P(self.H_final), P(gT), P(gfl), P(self.flags),
                   ctypes.addressof(c_last), P(lam), n_cells, st)
        seg_len = self.K
        for s_idx in reversed(range(len(self.ckpt))):
            s0 = s_idx * seg_len
            s1 = min(self.n_steps, s0 + seg_len)
            L = s1 - s0
            Hs = torch.empty((L,) + tuple(shp), dtype=torch.float32, device=dev)
            fls = torch.empty_like(Hs)
            As, Es, Fs = torch.empty_like(Hs), torch.empty_like(Hs), torch.empty_like(Hs)
            self._load(self.ckpt[s_idx])
            for t in range(L):
                Hs[t].copy_(tensor_of(self.H))
                fls[t].copy_(tensor_of(self.fl))
                self._step(s0 + t, self.theta, save={"A": As[t], "E": Es[t], "F": Fs[t]})
            for t in reversed(range(L)):
                n = s0 + t
                c = self._consts(n)
                if self.is_welding:
                    lib().call("waam_adj_arc_sums", P(self.flags), P(fls[t]), ctypes.addressof(c),
                               P(sums), g.nx, g.ny, g.nz, g.layout, P(g.scratch), st)
                lib().call("waam_adj_thermal_step", P(lam), P(lam2), P(Hs[t]), P(As[t]),
                           P(Es[t]), P(Fs[t]), P(fls[t]), P(self.flags), P(sums),
                           ctypes.addressof(c), P(grads), P(lamC), g.nx, g.ny, g.nz, g.layout,
                           P(g.scratch), st)
                lam, lam2 = lam2, lam
        self.lam0 = lam
        return grads.cpu().numpy()


class ThermalRolloutFn(torch.autograd.Function):
    """(T_final, f_l_final) = rollout(theta), differentiable in theta (a length-3 tensor)."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, rollout: ThermalRollout):
        T, fl = rollout.forward(theta.detach().double().cpu().numpy())
        ctx.rollout = rollout
        ctx.theta_dtype, ctx.theta_device = theta.dtype, theta.device
        return T, fl

    @staticmethod
    def backward(ctx, grad_T, grad_fl):
        g = ctx.rollout.backward(grad_T.float(), grad_fl.float())
        return torch.tensor(g, dtype=ctx.theta_dtype, device=ctx.theta_device), None


def field_loss(T, fl, T_target, fl_target, T_scale: float = 100.0):
    """Mean squared error of the final temperature (scaled) and liquid fraction fields."""
    return (((T - T_target) / T_scale) ** 2).mean() + ((fl - fl_target) ** 2).mean()


def fit_calibration(rollout: ThermalRollout, T_target, fl_target, theta0, iters: int = 20,
                    lr: float = 0.02) -> dict:
    """Gradient descent (Adam) on theta = (eta, s_h, s_sigma) against target fields; returns the
    parameter and loss history."""
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=lr)
    hist = {"theta": [], "loss": []}
    for _ in range(iters):
        opt.zero_grad()
        T, fl = ThermalRolloutFn.apply(theta, rollout)
        loss = field_loss(T, fl, T_target, fl_target)
        loss.backward()
        hist["theta"].append(theta.detach().numpy().copy())
        hist["loss"].append(float(loss))
        opt.step()
    return hist


# ---------------------------------------------------------------------------- Marangoni forcing
def marangoni_force(T: torch.Tensor, phi: torch.Tensor, f_l: torch.Tensor, flags: torch.Tensor,
                    dgamma: float, dims: tuple[int, int, int], layout: int):
    """The thermocapillary force (reference kernels.py:2158-2217) on device tensors, through the
    product kernel: (Fx, Fy, Fz), zero where it does not apply."""
    nx, ny, nz = dims
    F = [torch.zeros_like(T) for _ in range(3)]
    st = torch.cuda.current_stream().cuda_stream
    lib().call("waam_compute_marangoni_force", T.data_ptr(), phi.data_ptr(), f_l.data_ptr(),
               F[0].data_ptr(), F[1].data_ptr(), F[2].data_ptr(), flags.data_ptr(), float(dgamma),
               nx, ny, nz, layout, st)
    return tuple(F)


def marangoni_force_vjp(T, phi, f_l, flags, lam, dgamma: float, dims, layout: int):
    """Hand-written adjoint (``waam_adj_marangoni``) of ``marangoni_force`` with phi, f_l and
    the flags held: for lam = (dL/dFx, dL/dFy, dL/dFz), returns dL/dT (binary64, shaped like
    T) and dL/d(dgamma) (a binary64 sum over the cells)."""
    nx, ny, nz = dims
    lam = [x.contiguous().float() for x in lam]
    T_adj = torch.zeros(T.shape, dtype=torch.float64, device=T.device)
    s_cell = torch.empty(T.shape, dtype=torch.float64, device=T.device)
    G = torch.empty((3,) + tuple(T.shape), dtype=torch.float64, device=T.device)
    st = torch.cuda.current_stream().cuda_stream
    lib().call("waam_adj_marangoni", T.data_ptr(), phi.data_ptr(), f_l.data_ptr(),
               flags.data_ptr(), lam[0].data_ptr(), lam[1].data_ptr(), lam[2].data_ptr(),
               float(dgamma), T_adj.data_ptr(), s_cell.data_ptr(), G.data_ptr(), nx, ny, nz,
               layout, st)
    return T_adj, s_cell.sum()


class MarangoniForceFn(torch.autograd.Function):
    """(T, dgamma) -> (Fx, Fy, Fz) as a ``torch.autograd.Function``: forward through the
    product kernel, backward through the adjoint kernels (phi, f_l and the flags held, as the
    thermal rollout holds them)."""

    @staticmethod
    def forward(ctx, T, dgamma, phi, f_l, flags, dims, layout):
        ctx.save_for_backward(T, phi, f_l, flags)
        ctx.dgamma, ctx.dims, ctx.layout = float(dgamma), dims, layout
        return marangoni_force(T, phi, f_l, flags, float(dgamma), dims, layout)

    @staticmethod
    def backward(ctx, gx, gy, gz):
        T, phi, f_l, flags = ctx.saved_tensors
        zero = torch.zeros_like(T)
        lam = [g if g is not None else zero for g in (gx, gy, gz)]
        T_adj, d_dg = marangoni_force_vjp(T, phi, f_l, flags, lam, ctx.dgamma, ctx.dims,
                                          ctx.layout)
        g_T = T_adj.to(T.dtype) if ctx.needs_input_grad[0] else None
        g_dg = d_dg.reshape(()) if ctx.needs_input_grad[1] else None
        return g_T, g_dg, None, None, None, None, None


# ---------------------------------------------------------------------------- collision
def collide_srt_vjp(f_src, Fx, Fy, Fz, f_l, flags, lam, omega: float, C_darcy: float):
    """Hand-written adjoint (``waam_adj_collide_srt``) of the BGK collision with Guo forcing and
    the Carman-Kozeny drag (reference kernels.py:2815-2885), f_l and the flags held: for
    lam = dL/d(post-collision populations) (19 q-major n-cell arrays), returns dL/d(f_src)
    (float64, like f_src) and dL/d(Fx, Fy, Fz) (float64, (3,) + the field shape)."""
    n = f_l.numel()
    lam = lam.contiguous().float()
    lam_f = torch.empty(f_src.shape, dtype=torch.float64, device=f_src.device)
    lam_F = torch.empty((3,) + tuple(f_l.shape), dtype=torch.float64, device=f_src.device)
    st = torch.cuda.current_stream().cuda_stream
    lib().call("waam_adj_collide_srt", f_src.data_ptr(), Fx.data_ptr(), Fy.data_ptr(),
               Fz.data_ptr(), f_l.data_ptr(), flags.data_ptr(), lam.data_ptr(), float(omega),
               float(C_darcy), lam_f.data_ptr(), lam_F.data_ptr(), n, st)
    return lam_f, lam_F


# ---------------------------------------------------------------------------- the coupled step
class CoupledRollout:
    """n_steps of the coupled step with the flags and phi frozen, from the twin's current state
    along ``path`` (step -> torch (x_m, y_m)), differentiable in theta = (eta, s_h, s_sigma).

    Forward, per step (reference solvers/coupled_step.py:97-409 with the free surface, the flags,
    deposition and the stream held), the product's own kernels: arc injection and ceiling,
    conduction and enthalpy advection by the step's starting velocity, boundary losses, floor,
    phase map, ceiling, phase map; clear forces, CSF, Marangoni, arc pressure, hydrostatic
    gravity, buoyancy, force clamp; MRT collision and Mach clamp. With frozen flags the stream
    never feeds a collided cell (the next collision reads the cell's own post-collision
    populations), so the populations ping-pong between two buffers without it.

    Backward (``csrc/adjoint_coupled.hip``): the adjoint state (lambda_H, lambda_f, lambda_u) in
    binary64 runs back through each step — collision with the Mach clamp, forces with the force
    clamp, the phase map, the thermal chain with advection — and accumulates dL/dtheta in
    binary64. The forward keeps the state every ``checkpoint_every`` steps and the backward
    recomputes each segment from its checkpoint."""

    ADVECTION = ("flow", "frozen", "none")

    def __init__(self, twin, n_steps: int, path, checkpoint_every: int = 50,
                 is_welding: bool = True, advection: str = "flow"):
        """advection: "flow" — the enthalpy is advected by the velocity each step's collision
        produces (the coupled step); "frozen" — by the starting velocity throughout, with the
        forces and the collision left out; "none" — not at all (the thermal chain alone)."""
        if advection not in self.ADVECTION:
            raise ValueError(f"advection {advection!r} (one of {self.ADVECTION})")
        self.advection = advection
        if type(twin.arc_source).__name__ != "Gaussian2D":
            raise NotImplementedError("the differentiation form covers the Gaussian2D arc")
        if twin.use_srt or twin.use_material_tables:
            raise NotImplementedError("the coupled adjoint covers the scalar-material MRT step")
        if twin.enable_recoil or twin.enable_gas_shear or twin.enable_lorentz:
            raise NotImplementedError("the coupled adjoint covers the flow-tier forces")
        twin.sync()
        self.tw, g = twin, twin.grid
        self.n_steps, self.K = int(n_steps), max(1, int(checkpoint_every))
        self.path, self.is_welding = path, bool(is_welding)
        f = lambda: g.field(torch.float32)  # noqa: E731
        fp = lambda: g.field(torch.float32, leading=(19,))  # noqa: E731
        self.H, self.T, self.fl, self.Htmp = f(), f(), f(), f()
        self.rho, self.ux, self.uy, self.uz = f(), f(), f(), f()
        self.Fx, self.Fy, self.Fz = f(), f(), f()
        self.fa, self.fb = fp(), fp()
        self.flags = g.field(torch.uint8, np_dtype=np.int32)
        self.phi = f()
        tensor_of(self.flags).copy_(tensor_of(g.flags))
        tensor_of(self.phi).copy_(tensor_of(g.phi))
        # the starting state: the twin's enthalpy (T, f_l from the phase map), populations and
        # velocity
        self.H0 = tensor_of(g.H).clone()
        self.f0 = tensor_of(g.f_src).clone()
        self.u0 = [tensor_of(x).clone() for x in (g.ux, g.uy, g.uz)]
        self.zero_u = f()
        # the velocity the enthalpy is advected by when it does not evolve
        self.u_adv = None
        if advection == "frozen":
            self.u_adv = [f() for _ in range(3)]
            for d_, s_ in zip(self.u_adv, self.u0):
                tensor_of(d_).copy_(s_)
        elif advection == "none":
            self.u_adv = [self.zero_u] * 3
        self.h_conv0, self.eps0 = float(twin.h_conv), float(twin.eps_rad)
        self.sigma0 = float(twin.sigma_cells)
        x0, y0 = path(0)
        self.arc_k = resolve_arc_k(twin, g, x0 / g.dx, y0 / g.dx)
        self.norm_buf = torch.zeros((), dtype=torch.float32, device=g.device)
        self.u_cap = float(getattr(twin, "u_mach_limit_lu", 0.08))
        self.F_cap = float(getattr(twin, "force_limit_lu", 0.05))
        self.ckpt: list = []
        self.theta = None

    # ------------------------------------------------------------------ forward
    def _state_in(self, H, fpop, u) -> None:
        tensor_of(self.H).copy_(H)
        tensor_of(self.fa).copy_(fpop)
        for d, s in zip((self.ux, self.uy, self.uz), u):
            tensor_of(d).copy_(s)
        self._phase()

    def _state_out(self):
        return (tensor_of(self.H).clone(), tensor_of(self.fa).clone(),
                [tensor_of(x).clone() for x in (self.ux, self.uy, self.uz)])

    def _phase(self) -> None:
        tw = self.tw
        kernels.update_phase(self.H, self.T, self.fl, tw.cp_rho, tw.L_rho, tw.mat.T_solidus,
                             tw.mat.T_liquidus)

    def _ceiling(self) -> None:
        tw, g = self.tw, self.tw.grid
        if tw.enable_enthalpy_cap:
            kernels.clamp_enthalpy_ceiling_scalar(self.H, self.flags, tw.cp_rho,
                                                  tw.mat.T_solidus, tw.mat.T_liquidus,
                                                  tw.T_vapor_cap_K, tw.L_rho, g.FLAG_GAS)

    def _step(self, n: int, theta, save=None) -> None:
        tw, g = self.tw, self.tw.grid
        eta, s_h, s_sig = (float(v) for v in theta)
        x, y = self.path(n)
        arc_i, arc_j = x / g.dx, y / g.dx
        sv = ((lambda k, t: save[k].copy_(tensor_of(t)) if k in save else None)
              if save is not None else None)
        if sv:
            sv("H", self.H)
            sv("fl", self.fl)
            for a, u in zip(("ux", "uy", "uz"), (self.ux, self.uy, self.uz)):
                sv(a, u)
        if self.is_welding:
            kernels.inject_arc_heat(self.H, self.flags, self.phi, self.fl, self.norm_buf, arc_i,
                                    arc_j, self.arc_k, tw.Q_w, self.sigma0 * s_sig, g.dt,
                                    g.dx ** 3, eta, tw.arc_penetration_m / g.dx,
                                    1 if tw.arc_surface_weighting else 0, g.FLAG_SOLID,
                                    g.FLAG_GAS, scratch=g.scratch)
        if sv:
            sv("A", self.H)
        self._ceiling()
        if sv:
            sv("B", self.H)
        ua = self.u_adv or (self.ux, self.uy, self.uz)
        kernels.advect_diffuse_temperature_into(self.H, self.Htmp, self.T, ua[0], ua[1], ua[2],
                                                self.flags, tw.alpha_lu, 1.0, tw.cp_rho, g.nx,
                                                g.ny, g.nz)
        self.H, self.Htmp = self.Htmp, self.H
        if tw.enable_heat_loss:
            kernels.apply_thermal_boundary_losses(
                self.H, self.T, self.flags, tw.T_amb, self.h_conv0 * s_h, self.eps0 * s_h,
                1 if tw.enable_convection else 0, 1 if tw.enable_radiation else 0, tw.cp_rho,
                g.dt, g.dx, tw.sigma_sb, g.FLAG_SOLID, g.FLAG_GAS, g.nx, g.ny, g.nz)
        if sv:
            sv("E", self.H)
        if tw.enable_heat_loss:
            kernels.clamp_enthalpy_floor_scalar(self.H, tw.cp_rho, self.flags, tw.T_amb,
                                                g.FLAG_GAS)
        if sv:
            sv("F", self.H)
        self._phase()
        self._ceiling()
        self._phase()
        if self.advection != "flow":
            return
        # forces on the new temperature and liquid fraction
        kernels.clear_forces(self.Fx, self.Fy, self.Fz)
        if tw.enable_csf_tension:
            kernels.compute_csf_tension(self.phi, self.flags, self.Fx, self.Fy, self.Fz,
                                        tw.gamma_lu, g.FLAG_SOLID, g.FLAG_GAS, g.nx, g.ny, g.nz,
                                        int(tw.enable_wetting), tw.theta_rad)
        kernels.compute_marangoni_force(self.T, self.phi, self.fl, self.Fx, self.Fy, self.Fz,
                                        self.flags, tw.dgamma_dT_lu, g.dx, g.FLAG_SOLID,
                                        g.FLAG_GAS, g.nx, g.ny, g.nz)
        if self.is_welding:
            kernels.apply_arc_pressure(self.Fz, self.flags, self.phi, arc_i, arc_j, self.arc_k,
                                       weld_forces.arc_pressure_sigma_cells(tw),
                                       weld_forces.arc_pressure_peak_pa(tw), g.dt, g.dx,
                                       tw.mat.rho, g.FLAG_SOLID, g.FLAG_GAS)
        if tw.enable_hydrostatic_gravity:
            kernels.add_hydrostatic_gravity(self.Fz, self.fl, self.flags, 1.0, tw.g_lu,
                                            g.FLAG_SOLID, g.FLAG_GAS)
        kernels.add_buoyancy(self.T, self.Fz, self.fl, self.flags, tw.g_lu, tw.beta_T,
                             tw.mat.T_liquidus, 1.0, g.FLAG_SOLID, g.FLAG_GAS)
        if sv:
            for a, fld in zip(("Frx", "Fry", "Frz"), (self.Fx, self.Fy, self.Fz)):
                sv(a, fld)
        if self.F_cap > 0.0:
            kernels.clamp_body_force_magnitude(self.Fx, self.Fy, self.Fz, self.flags, self.F_cap,
                                               g.FLAG_SOLID, g.FLAG_GAS)
        if sv:
            for a, fld in zip(("Fx", "Fy", "Fz"), (self.Fx, self.Fy, self.Fz)):
                sv(a, fld)
            sv("f", self.fa)
            sv("H1", self.H)
            sv("T1", self.T)
            sv("fl1", self.fl)
        kernels.collide_mrt(self.fa, self.fb, self.rho, self.ux, self.uy, self.uz, self.Fx,
                            self.Fy, self.Fz, self.fl, self.flags, None, None, None, None, None,
                            tw.omega, tw.omega_bulk, tw.C_darcy, g.FLAG_SOLID, g.FLAG_GAS, g.nx,
                            g.ny, g.nz)
        if self.u_cap > 0.0:
            kernels.clamp_velocity_mach(self.ux, self.uy, self.uz, self.flags, self.u_cap,
                                        g.FLAG_SOLID, g.FLAG_GAS)
        self.fa, self.fb = self.fb, self.fa

    def forward(self, theta) -> tuple[torch.Tensor, torch.Tensor]:
        """Final (T, f_l) (storage order) after n_steps with parameters theta."""
        self.theta = tuple(float(v) for v in theta)
        self.ckpt = []
        self._state_in(self.H0, self.f0, self.u0)
        for n in range(self.n_steps):
            if n % self.K == 0:
                self.ckpt.append(self._state_out())
            self._step(n, self.theta)
        self.H_final = tensor_of(self.H).clone()
        return tensor_of(self.T).clone(), tensor_of(self.fl).clone()

    # ------------------------------------------------------------------ backward
    _SAVE = ("H", "fl", "ux", "uy", "uz", "A", "B", "E", "F", "Frx", "Fry", "Frz", "Fx", "Fy",
             "Fz", "H1", "T1", "fl1")

    def backward(self, grad_T: torch.Tensor, grad_fl: torch.Tensor) -> np.ndarray:
        """dL/dtheta (float64) given dL/dT and dL/df_l at the final step."""
        if self.theta is None:
            raise RuntimeError("backward before forward")
        tw, g = self.tw, self.tw.grid
        st = torch.cuda.current_stream().cuda_stream
        dev = tensor_of(self.H).device
        N = g.n_cells
        shp = tuple(tensor_of(self.H).shape)
        d64 = dict(dtype=torch.float64, device=dev)
        flow = self.advection == "flow"
        lam_H = torch.zeros(shp, **d64)
        lam_Ho = torch.empty(shp, **d64)
        lam_u = torch.zeros((3,) + shp, **d64)
        lam_uo = torch.empty((3,) + shp, **d64)
        lamC = torch.empty(shp, **d64)
        lamB = torch.empty(shp, **d64)
        grads = torch.zeros(3, **d64)
        sums = torch.zeros(2, **d64)
        if flow:
            lam_f = torch.zeros((19,) + shp, **d64)
            lam_fin = torch.empty((19,) + shp, **d64)
            lam_F = torch.empty((3,) + shp, **d64)
            lam_Fr = torch.empty((3,) + shp, **d64)
            lam_T1 = torch.empty(shp, **d64)
            lam_fl1 = torch.empty(shp, **d64)
            G = torch.empty((3,) + shp, **d64)
        c_last = self._consts(self.n_steps - 1)
        # the binary64 seeds live in locals while their kernel is queued (P() keeps no reference)
        gT = grad_T.double().contiguous()
        gfl = grad_fl.double().contiguous()
        lib().call("waam_cadj_phase", P(self.H_final), P(self.flags), ctypes.addressof(c_last),
                   P(gT), P(gfl), P(lam_H), N, st)
        keys = self._SAVE if flow else ("H", "fl", "A", "B", "E", "F")
        for s_idx in reversed(range(len(self.ckpt))):
            s0 = s_idx * self.K
            s1 = min(self.n_steps, s0 + self.K)
            L = s1 - s0
            sv = {k: torch.empty((L,) + shp, dtype=torch.float32, device=dev) for k in keys}
            if flow:
                sv["f"] = torch.empty((L, 19) + shp, dtype=torch.float32, device=dev)
            self._state_in(*self.ckpt[s_idx])
            for t in range(L):
                self._step(s0 + t, self.theta, save={k: v[t] for k, v in sv.items()})
            for t in reversed(range(L)):
                c = self._consts(s0 + t)
                if flow:
                    # collision (with the Mach clamp) -> lam_f of the step's populations,
                    # lam_F and f_l
                    lam_fl1.zero_()
                    lam_T1.zero_()
                    lib().call("waam_cadj_collide_mrt", P(sv["f"][t]), P(sv["Fx"][t]),
                               P(sv["Fy"][t]), P(sv["Fz"][t]), P(sv["fl1"][t]), P(self.flags),
                               P(lam_f), P(lam_u), tw.omega, tw.omega_bulk, tw.C_darcy,
                               self.u_cap, P(lam_fin), P(lam_F), P(lam_fl1), N, st)
                    lib().call("waam_cadj_forces", P(sv["Frx"][t]), P(sv["Fry"][t]),
                               P(sv["Frz"][t]), P(sv["T1"][t]), P(sv["fl1"][t]), P(self.flags),
                               P(lam_F), self.F_cap, tw.g_lu, tw.beta_T, tw.mat.T_liquidus,
                               1 if tw.enable_hydrostatic_gravity else 0, P(lam_Fr), P(lam_T1),
                               P(lam_fl1), N, st)
                    lib().call("waam_cadj_marangoni", P(self.phi), P(sv["fl1"][t]),
                               P(self.flags), P(lam_Fr), tw.dgamma_dT_lu, P(G), P(lam_T1), g.nx,
                               g.ny, g.nz, g.layout, st)
                    lib().call("waam_cadj_phase", P(sv["H1"][t]), P(self.flags),
                               ctypes.addressof(c), P(lam_T1), P(lam_fl1), P(lam_H), N, st)
                    ua = (sv["ux"][t], sv["uy"][t], sv["uz"][t])
                else:
                    ua = tuple(tensor_of(x) for x in self.u_adv)
                # the thermal chain with advection -> lam_H and lam_u of the step's start
                if self.is_welding:
                    lib().call("waam_adj_arc_sums", P(self.flags), P(sv["fl"][t]),
                               ctypes.addressof(c), P(sums), g.nx, g.ny, g.nz, g.layout,
                               P(g.scratch), st)
                lib().call("waam_cadj_thermal_step", P(lam_H), P(lam_Ho), P(lam_uo),
                           P(sv["H"][t]), P(sv["A"][t]), P(sv["B"][t]), P(sv["E"][t]),
                           P(sv["F"][t]), P(sv["fl"][t]), P(ua[0]), P(ua[1]), P(ua[2]),
                           P(self.flags), P(sums), ctypes.addressof(c), P(grads), P(lamC),
                           P(lamB), g.nx, g.ny, g.nz, g.layout, P(g.scratch), st)
                lam_H, lam_Ho = lam_Ho, lam_H
                if flow:
                    lam_u, lam_uo = lam_uo, lam_u
                    lam_f, lam_fin = lam_fin, lam_f
            del sv
        self.lam0 = lam_H
        return grads.cpu().numpy()

    def _consts(self, n: int) -> AdjConsts:
        return ThermalRollout._consts(self, n)


class CoupledRolloutFn(torch.autograd.Function):
    """(T_final, f_l_final) = coupled rollout(theta), differentiable in theta (length 3)."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, rollout: CoupledRollout):
        T, fl = rollout.forward(theta.detach().double().cpu().numpy())
        ctx.rollout = rollout
        ctx.theta_dtype, ctx.theta_device = theta.dtype, theta.device
        return T, fl

    @staticmethod
    def backward(ctx, grad_T, grad_fl):
        g = ctx.rollout.backward(grad_T, grad_fl)
        return torch.tensor(g, dtype=ctx.theta_dtype, device=ctx.theta_device), None


def pool_wd_surrogate(fl: torch.Tensor, i_c: int, nz_solid: int, layout: int, dx: float,
                      tau: float = 0.1) -> tuple[torch.Tensor, torch.Tensor]:
    """Smoothed pool width and depth (m) on the cross-section x = i_c, differentiable in f_l.

    The reference's telemetry (twin.py:716-744) counts, on the cross-section through the pool,
    the y-extent of f_l > 0.5 and the plate layers down to the deepest f_l > 0.5: step functions
    of f_l. The surrogate replaces each column's and each plate layer's indicator with a logistic
    of its largest liquid fraction: W = dx * sum_j s((max_k f_l(i_c, j, k) - 0.5) / tau),
    D = dx * sum_{k < nz_solid} s((max_j f_l(i_c, j, k) - 0.5) / tau), s the logistic function;
    as tau -> 0 these become the cell counts of a pool without holes."""
    sec = fl[:, :, i_c] if layout == 1 else fl[i_c].transpose(0, 1)  # (nz, ny)
    col = sec.max(dim=0).values            # per j: the column's largest f_l
    lay = sec[:nz_solid].max(dim=1).values  # per plate layer
    W = dx * torch.sigmoid((col - 0.5) / tau).sum()
    D = dx * torch.sigmoid((lay - 0.5) / tau).sum()
    return W, D
