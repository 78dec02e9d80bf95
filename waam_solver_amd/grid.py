"""WAAMGrid — the coupled step's state as device tensors (reference grid.py:37-289).

Every field the reference allocates with ``ti.field`` is a contiguous torch tensor on the ROCm
device, wrapped in a DeviceField, with the same names; populations are 19 such arrays. Flags
are uint8. Lattice-unit scaling and the stability checks follow reference grid.py:89-126.

Storage layout (``layout``): ``LAYOUT_ZYX`` (default) stores every field as (nz, ny, nx), i
fastest, so the GPU's wavefronts run along horizontal rows, which the plate, the free surface
and the gas above fill whole; ``LAYOUT_XYZ`` is the reference's (nx, ny, nz), k fastest. The
fields' accessors (``to_numpy``, ``from_numpy``, ``[i, j, k]``) present (nx, ny, nz) either way.

Memory: 2 x 19 x 4 B of populations plus about 26 scalar fields of 4 B per cell — 4.3 GB of
populations and ~5.9 GB of fields at 768x384x192 — well inside one MI355X's 288 GB.
"""

from __future__ import annotations

import numpy as np
import torch

from .field import LAYOUT_XYZ, LAYOUT_ZYX, DeviceField
from .materials import MaterialProps

Q = 19
EX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
EY = (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1)
EZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1)
W = (1.0 / 3.0,) + (1.0 / 18.0,) * 6 + (1.0 / 36.0,) * 12
OPP = (0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17)


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("waam_solver_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


class WAAMGrid:
    FLAG_FLUID = 0
    FLAG_SOLID = 1
    FLAG_GAS = 2
    FLAG_IFACE = 4

    def __init__(self, nx: int, ny: int, nz: int, dx: float, mat: MaterialProps,
                 max_tracers: int = 50_000, allocate_lorentz: bool = False,
                 allocate_vof: bool = False, allocate_export: bool = False,
                 device: torch.device | str | None = None, layout: int = LAYOUT_ZYX):
        if layout not in (LAYOUT_XYZ, LAYOUT_ZYX):
            raise ValueError(f"unknown field layout {layout}")
        self.layout = int(layout)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = float(dx)
        self.mat = mat
        self.max_tracers = int(max_tracers)
        self.device = torch.device(device) if device is not None else default_device()

        # lattice units (reference grid.py:93-105)
        self.u_ref_phys = 0.5
        self.u_ref_lu = 0.05
        self.dt = dx * self.u_ref_lu / self.u_ref_phys
        nu_phys = mat.mu / mat.rho
        nu_lu = nu_phys * self.dt / (dx ** 2)
        self.tau = 3.0 * nu_lu + 0.5
        alpha_lu = mat.alpha * self.dt / (dx ** 2)
        self.alpha_lu = alpha_lu
        self.tau_T = 3.0 * alpha_lu + 0.5
        if self.tau <= 0.5:
            raise ValueError(f"LBM unstable: tau={self.tau:.4f} <= 0.5 (nu_lu={nu_lu:.3e}). "
                             "Reduce dx or increase viscosity.")
        if alpha_lu > 1.0 / 6.0:
            raise ValueError(f"Thermal step unstable: alpha_lu={alpha_lu:.4f} > 1/6 "
                             "(explicit 3D diffusion limit). Reduce dx or dt ratio.")

        s3 = (self.nx, self.ny, self.nz)
        self._shape3 = s3
        dev = self.device
        f32 = lambda: self.field(torch.float32)  # noqa: E731
        # populations: f_b starts at zero, as a freshly allocated Taichi field does
        self.f_a = self.field(torch.float32, leading=(Q,))
        self.f_b = self.field(torch.float32, leading=(Q,))
        for name in ("rho", "T", "T_max", "H", "f_l", "phi", "cp_rho_field", "alpha_lu_field",
                     "dgamma_lu_field", "tau_field", "dT_dt", "T_prev", "time_above_800_s",
                     "time_above_1100_s", "time_above_solidus_s", "Fx_snap", "Fy_snap",
                     "Fz_snap", "ux", "uy", "uz", "Fx", "Fy", "Fz"):
            setattr(self, name, f32())
        self.flags = self.field(torch.uint8, np_dtype=np.int32)
        # Jacobi partner of H (thermal advection reads the pre-step enthalpy)
        self.H_tmp = f32()
        # 0-d buffers of the reference (grid.py:160-172)
        z0 = lambda dt=torch.float32: DeviceField.zeros((), dt, dev)  # noqa: E731
        self.surface_k_buf = z0()
        self.deposit_vol_buf = z0()
        self.deposit_real_buf = z0()
        self.arc_norm_buf = z0()
        self.deposit_result = DeviceField.zeros((2,), torch.float32, dev)
        # the run's deposition ledger, kept on the device: [0] placed volume (m^3, binary64 sum
        # in droplet order), [1] droplets that placed less than 98 % of their volume
        self.deposit_ledger = DeviceField.zeros((2,), torch.float64, dev)
        self.telem_out = DeviceField.zeros((8,), torch.float64, dev)
        self.bead_top_buf = DeviceField.zeros((1,), torch.int32, dev)
        self.telem_fl_slice = DeviceField.zeros((self.ny, self.nz), torch.float32, dev)
        # tracers
        self.porosity_pos = DeviceField.zeros((self.max_tracers, 3), torch.float32, dev)
        self.porosity_active = DeviceField.zeros((self.max_tracers,), torch.int32, dev)
        self.tracer_head = DeviceField.zeros((), torch.int32, dev)
        # device workspace for reductions and deposition (include/waam_hip.h)
        from ._lib import lib
        self.scratch = torch.zeros(lib().scratch_bytes(), dtype=torch.uint8, device=dev)

        self._has_lorentz = self._has_vof_buf = self._has_export = False
        self._n_core_volume = 24
        self._n_optional_volume = 0
        if allocate_vof:
            self.ensure_vof_buffers()
        if allocate_lorentz:
            self.ensure_lorentz_fields()
        if allocate_export:
            self.ensure_export_buffers()
        self._read_buf = self.f_a
        self._write_buf = self.f_b

    def field(self, dtype, np_dtype=None, leading: tuple = ()) -> DeviceField:
        """New zero grid field in this grid's layout (logical shape leading + (nx, ny, nz))."""
        return DeviceField.zeros(tuple(leading) + self._shape3, dtype, self.device,
                                 np_dtype=np_dtype, layout=self.layout)

    # ------------------------------------------------------------------ optional buffers
    def ensure_vof_buffers(self) -> None:
        if self._has_vof_buf:
            return
        self.phi_tmp = self.field(torch.float32)
        self._has_vof_buf = True
        self._n_optional_volume += 1

    def ensure_lorentz_fields(self) -> None:
        """Potential (and its Jacobi partner), conductivity, current source, J, B, the radial
        current bins (nz x 64) and the two L1 sums (reference grid.py:235-254)."""
        if getattr(self, "_has_lorentz", False):
            return
        for name in ("phi_elec", "phi_elec_tmp", "sigma_elec", "elec_source", "Jx", "Jy", "Jz",
                     "Bx", "By", "Bz"):
            setattr(self, name, self.field(torch.float32))
        self.n_elec_bins = 64
        self.elec_rad_bins = DeviceField.zeros((self.nz, self.n_elec_bins), torch.float32,
                                               self.device)
        self.elec_l1_buf = DeviceField.zeros((2,), torch.float32, self.device)
        self._has_lorentz = True
        self._n_optional_volume += 10

    def ensure_export_buffers(self) -> None:
        if self._has_export:
            return
        self.kappa_field = self.field(torch.float32)
        self.vorticity_mag = self.field(torch.float32)
        self._has_export = True
        self._n_optional_volume += 2

    # ------------------------------------------------------------------ buffers
    def swap_buffers(self) -> None:
        self._read_buf, self._write_buf = self._write_buf, self._read_buf

    @property
    def f_src(self) -> DeviceField:
        return self._read_buf

    @property
    def f_dst(self) -> DeviceField:
        return self._write_buf

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def estimated_vram_mb(self) -> float:
        n = self.n_cells
        dist = 2 * Q * n * 4
        n_vol = self._n_core_volume + self._n_optional_volume + 2
        return (dist + n_vol * n * 4 + self.ny * self.nz * 4 + self.max_tracers * 16) / 1024 ** 2
