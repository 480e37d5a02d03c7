"""Synthetic inputs for the v-Poisson / derivative path.

Grid and wavepacket right-hand sides used by the tests and the benchmark.
The grid follows the reference's geometry: uniform x with spacing ``dx``,
geometric y with first spacing ``dy0`` and ratio ``stf`` (main_par.f:1841-1870),
``nmodes`` spanwise Fourier modes with kz = k*beta, k = 0..nmodes-1
(main_par.f:1886-1889).
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np


@dataclasses.dataclass(frozen=True)
class Grid:
    nx: int
    ny: int
    nmodes: int
    levels: int = 4
    dx: float = 0.02
    dy0: float = 5e-4
    stf: float = 1.02
    alpha: float = 0.0
    beta: float = 0.2
    fac_y: float = 1.0

    def x(self):
        return np.arange(self.nx) * self.dx

    def y(self):
        j = np.arange(self.ny)
        if self.stf == 1.0:
            return self.dy0 * j
        return self.dy0 * (self.stf ** j - 1.0) / (self.stf - 1.0)

    def level_shape(self, lvl):
        nx, ny = self.nx, self.ny
        for _ in range(lvl - 1):
            nx, ny = (nx + 1) // 2, (ny + 1) // 2
        return nx, ny


# Configurations named by their sizes (the five workloads of this path)
SMALL_385x97x8 = Grid(385, 97, 8, 4, 0.02, 5e-4, 1.02, 0.0, 0.2)
BATCH_1025x193x64 = Grid(1025, 193, 64, 4, 0.02, 2e-4, 1.015, 0.0, 0.1)
LARGE_2049x257x128 = Grid(2049, 257, 128, 4, 0.02, 2e-4, 1.012, 0.0, 0.05)


def _seed(tag, problem, mode, part):
    h = hashlib.sha256(f"{tag}/{problem}/{mode}/{part}".encode()).digest()
    return int.from_bytes(h[:8], "little")


def wavepacket_rhs(grid: Grid, tag: str = "rhs", problem: int = 0, npackets: int = 6,
                   shift: float = 0.0):
    """Complex RHS f of shape (nmodes, ny, nx) (C view of Fortran (nx,ny,K)).

    Each mode and each of (re, im) is a sum of ``npackets`` Gaussian
    wavepackets A*exp(-((x-xc)^2+(y-yc)^2)/(2 s^2))*cos(a*x) with A~N(0,1),
    s~U[0.05,0.3], a~U[10,30], xc~U[0.1Lx,0.9Lx] (+shift, for packets
    drifting downstream), yc~U[0,0.5*ytop]; PCG64 seeded from
    (tag, problem, mode, part).
    """
    x = grid.x()
    y = grid.y()
    lx, ytop = x[-1], y[-1]
    X = x[None, :]
    Y = y[:, None]
    out = np.zeros((grid.nmodes, grid.ny, grid.nx), dtype=np.complex128)
    for k in range(grid.nmodes):
        for part in range(2):
            rng = np.random.Generator(np.random.PCG64(_seed(tag, problem, k, part)))
            acc = np.zeros((grid.ny, grid.nx))
            for _ in range(npackets):
                amp = rng.normal()
                s = rng.uniform(0.05, 0.3)
                a = rng.uniform(10.0, 30.0)
                xc = rng.uniform(0.1 * lx, 0.9 * lx) + shift
                yc = rng.uniform(0.0, 0.5 * ytop)
                acc += amp * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2 * s * s)) \
                    * np.cos(a * X)
            if part == 0:
                out[k].real = acc
            else:
                out[k].imag = acc
    return out
