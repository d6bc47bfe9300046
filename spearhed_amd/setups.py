"""Problem set-ups as lattice blocks for sph_init_run (initial conditions generated on the GPU).

Each function returns (run config, blocks, particle count) for the same physical set-up as the
host generator of the same name in ic.py (there with a host random jitter, here with the device
generator), so a serving request can carry a few hundred bytes of blocks instead of its particles.
"""
from __future__ import annotations

import math

from .capi import (SETUP_ORSZAG_TANG, SETUP_ROTOR, SETUP_SHEAR_LAYER, SETUP_UNIFORM, make_block,
                   make_config)
from .ic import KERNEL_WENDLAND, h_for_spacing


def brio_wu(n_target: int, kernel: int = KERNEL_WENDLAND):
    """Brio & Wu tube in the periodic box [0,1] x [0,w]^2, w = 16 dx: left half dx lattice (rho 1,
    P 1, By +1), right half 2 dx lattice (rho 0.125, P 0.1, By -1), Bx 0.75, gamma 2, equal masses."""
    gamma = 2.0
    nyz = 16
    nxl = max(8, int(round(n_target / (nyz * nyz * (1 + 1.0 / 16)))))
    nxl -= nxl % 2
    dx = 0.5 / nxl
    w = nyz * dx
    h = h_for_spacing(dx, kernel)
    blocks = [make_block((0, 0, 0), (0.5, w, w), dx, 1.0, 1.0, bfield=(0.75, 1.0, 0.0)),
              make_block((0.5, 0, 0), (1.0, w, w), 2 * dx, 0.125, 0.1, bfield=(0.75, -1.0, 0.0))]
    cfg = make_config(box_len=(1.0, w, w), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=gamma, mhd=1,
                      cleaning=1, borve=1)
    return cfg, blocks, nxl * nyz * nyz + (nxl // 2) * (nyz // 2) ** 2


def rotor(n_target: int, kernel: int = KERNEL_WENDLAND):
    """Magnetised rotor in the periodic slab [0,1]^2 x [0, 4 dx]: disc r < 0.1 of density 10 in rigid
    rotation with edge speed 2, ambient density 1 at rest, P 1, Bx = 5 / sqrt(4 pi), gamma 1.4."""
    n = max(16, int(round(math.sqrt(n_target / 4))))
    dx = 1.0 / n
    h = h_for_spacing(dx, kernel)
    blocks = [make_block((0, 0, 0), (1.0, 1.0, 4 * dx), dx, 1.0, 1.0, bfield=(5.0 / math.sqrt(4 * math.pi), 0, 0),
                         setup=SETUP_ROTOR, params=(0.5, 0.5, 0.1, 2.0, 10.0))]
    cfg = make_config(box_len=(1.0, 1.0, 4 * dx), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=1.4, mhd=1,
                      cleaning=1, borve=1)
    return cfg, blocks, n * n * 4


def orszag_tang(n_target: int, kernel: int = KERNEL_WENDLAND):
    """Orszag-Tang vortex on [0,1]^2 x [0, 4 dx]: rho 25/(36 pi), P 5/(12 pi), v0 = 1,
    B0 = 1/sqrt(4 pi), gamma 5/3."""
    n = max(16, int(round(math.sqrt(n_target / 4))))
    dx = 1.0 / n
    h = h_for_spacing(dx, kernel)
    blocks = [make_block((0, 0, 0), (1.0, 1.0, 4 * dx), dx, 25.0 / (36.0 * math.pi), 5.0 / (12.0 * math.pi),
                         setup=SETUP_ORSZAG_TANG, params=(1.0, 1.0 / math.sqrt(4.0 * math.pi)))]
    cfg = make_config(box_len=(1.0, 1.0, 4 * dx), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=5.0 / 3.0,
                      mhd=1, cleaning=1, borve=1)
    return cfg, blocks, n * n * 4


def kelvin_helmholtz(n_target: int, dV: float = 1.0, beta: float = 100.0, phase: float = 0.0,
                     kernel: int = KERNEL_WENDLAND):
    """Shear layer on [0,1]^2 x [0, nz dx]: density 2 / 1 inside / outside |y - 1/2| < 1/4, vx = +-dV/2,
    vy = 0.01 sin(4 pi x + phase), P 2.5, Bx from the plasma beta, gamma 5/3."""
    n = max(8, int(round(n_target ** (1.0 / 3.0))))
    nz = max(3, int(round(n_target / (n * n))))
    dx = 1.0 / n
    h = h_for_spacing(dx, kernel)
    P = 2.5
    blocks = [make_block((0, 0, 0), (1.0, 1.0, nz * dx), dx, 1.0, P, bfield=(math.sqrt(2 * P / beta), 0, 0),
                         setup=SETUP_SHEAR_LAYER, params=(dV, 2.0, 0.01, phase))]
    cfg = make_config(box_len=(1.0, 1.0, nz * dx), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=5.0 / 3.0,
                      mhd=1, cleaning=1, borve=1)
    return cfg, blocks, n * n * nz


def uniform_box(n: int, density: float = 1.0, pressure: float = 1.0, kernel: int = KERNEL_WENDLAND):
    """A uniform periodic unit cube on an n^3 lattice at rest."""
    dx = 1.0 / n
    blocks = [make_block((0, 0, 0), (1, 1, 1), dx, density, pressure, setup=SETUP_UNIFORM)]
    cfg = make_config(periodic=(1, 1, 1), kernel=kernel, h0=h_for_spacing(dx, kernel))
    return cfg, blocks, n ** 3


DEVICE_SERVING_PROBLEMS = {"brio_wu": brio_wu, "rotor": rotor, "orszag_tang": orszag_tang}
