"""Lid-driven cavity: the plain D3Q19 central-moment MRT flow on its own.

A cubic box of n^3 fluid cells with no-slip walls on five faces and a lid moving at ``u_lid``
(lattice units) along x at the top (+z) face. The walls are a one-cell layer around the fluid
(arrays of (n + 2)^3 cells) with half-way bounce-back, so they sit half a cell outside the
outermost fluid cells and the fluid spans L = n cells. The viscosity follows from the Reynolds number,
nu = u_lid L / Re, tau = 3 nu + 1/2, both MRT rates 1/tau. The state starts at rest,
f = w_q rho, with rho = 1 + U(-noise, noise) from a seeded generator.

The coupled step's collision map is reused as is (``collide_mrt``); the stream is a physical
pull (every fluid cell, every step) with Ladd's moving-wall bounce-back, which the coupled
step's stream does not have (``waam_cavity_step``).
"""

from __future__ import annotations

import numpy as np
import torch

from . import kernels
from .field import LAYOUT_ZYX, DeviceField
from .grid import Q, W, default_device

FLUID, WALL, LID = 0, 1, 3


def cavity_flags(nx: int, ny: int, nz: int) -> np.ndarray:
    """(nx, ny, nz) flags: walls on the x, y faces and the bottom, the lid on the whole top."""
    fl = np.full((nx, ny, nz), WALL, np.uint8)
    fl[1:-1, 1:-1, 1:-1] = FLUID
    fl[:, :, -1] = LID
    return fl


def initial_populations(nx: int, ny: int, nz: int, noise: float, seed: int) -> tuple:
    """(f (19, nx, ny, nz), rho): rest equilibrium f = w_q rho, rho = 1 + U(-noise, noise)."""
    rng = np.random.default_rng(seed)
    rho = (1.0 + rng.uniform(-noise, noise, (nx, ny, nz))).astype(np.float32)
    w = np.array(W, np.float32)
    f = (w[:, None, None, None] * rho[None]).astype(np.float32)
    return f, rho


class LidDrivenCavity:
    def __init__(self, n: int = 64, u_lid: float = 0.05, reynolds: float = 100.0,
                 noise: float = 1e-4, seed: int = 0, device=None, layout: int = LAYOUT_ZYX):
        self.n = int(n)
        self.nx = self.ny = self.nz = self.n + 2
        self.L = self.n
        self.u_lid = float(u_lid)
        self.reynolds = float(reynolds)
        self.nu = self.u_lid * self.L / self.reynolds
        self.tau = 3.0 * self.nu + 0.5
        self.omega = 1.0 / self.tau
        self.device = torch.device(device) if device is not None else default_device()
        s3 = (self.nx, self.ny, self.nz)
        z = lambda dt=torch.float32, lead=(): DeviceField.zeros(  # noqa: E731
            lead + s3, dt, self.device, layout=layout)
        self.f_a, self.f_b = z(lead=(Q,)), z(lead=(Q,))
        self.rho, self.ux, self.uy, self.uz = z(), z(), z(), z()
        self.flags = DeviceField.zeros(s3, torch.uint8, self.device, np_dtype=np.int32,
                                       layout=layout)
        self.flags.from_numpy(cavity_flags(*s3))
        f, rho = initial_populations(*s3, noise, seed)
        self.f_a.from_numpy(f)
        self.f_b.from_numpy(f)
        self.rho.from_numpy(rho)
        self.steps = 0

    @property
    def f(self) -> DeviceField:
        """The current populations (post-collision of the last step)."""
        return self.f_a if self.steps % 2 == 0 else self.f_b

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            src, dst = (self.f_a, self.f_b) if self.steps % 2 == 0 else (self.f_b, self.f_a)
            kernels.cavity_step(src, dst, self.rho, self.ux, self.uy, self.uz, self.flags,
                                self.omega, self.omega, (self.u_lid, 0.0, 0.0), self.nx,
                                self.ny, self.nz)
            self.steps += 1

    def velocity(self) -> np.ndarray:
        """(3, nx, ny, nz) lattice velocity (zero in wall cells)."""
        return np.stack([self.ux.to_numpy(), self.uy.to_numpy(), self.uz.to_numpy()])

    def centreline_profiles(self) -> dict[str, np.ndarray]:
        """u_x / u_lid along the vertical centreline and u_z / u_lid along the horizontal one
        (mid-plane y), at the fluid cells, with their positions in units of L from the bottom
        wall / the x = 0 wall (the walls sit half a cell outside the outer fluid cells)."""
        ux, uz = self.ux.to_numpy(), self.uz.to_numpy()
        jm = self.ny // 2
        im, km = self.nx // 2, self.nz // 2
        zs = (np.arange(1, self.nz - 1) - 0.5) / self.L
        xs = (np.arange(1, self.nx - 1) - 0.5) / self.L
        return {"z": zs, "ux_vertical": ux[im, jm, 1:-1] / self.u_lid,
                "x": xs, "uz_horizontal": uz[1:-1, jm, km] / self.u_lid}

    def mass(self) -> float:
        """Sum of the populations of the fluid cells, in float64."""
        f = self.f.to_numpy().astype(np.float64)
        fluid = self.flags.to_numpy() == FLUID
        return float(f[:, fluid].sum())
