"""Seeded initial conditions for the benchmark and test problems.

All problems are jittered cubic lattices: a particle sits at (i + 1/2) dx + U(-0.05, 0.05) dx per
axis, drawn from numpy's Philox generator keyed by the seed. Masses are rho dx^3, so density
contrasts come either from lattice spacing (Sod, Brio-Wu: equal masses) or from mass (Kelvin-
Helmholtz, rotor). The smoothing length is constant per run, as in the reference
(include/spearhed/param/sph.param h0), set for ~57 neighbours on the densest lattice:
H = support * h = (3 * 57 / (4 pi))^(1/3) dx.

The reference's own Sod set-up (share/spearhed/ValidationSetups/SodShockTube.hpp:116-176: equal mass
rho_L dx_L^3, u = P / ((gamma - 1) rho), frozen wall slabs) is the model for sod_tube().
"""
from __future__ import annotations

import math

import numpy as np

from .capi import (KERNEL_QUINTIC, KERNEL_WENDLAND, SPECIES_BOUNDARY, SPECIES_DEFAULT, RunConfig,
                   make_config)
from .particles import ParticleSet

N_NEIGHBOURS = 57


def h_for_spacing(dx: float, kernel: int = KERNEL_WENDLAND, nb: int = N_NEIGHBOURS) -> float:
    support = 3 if kernel == KERNEL_QUINTIC else 2
    H = (3.0 * nb / (4.0 * math.pi)) ** (1.0 / 3.0) * dx
    return H / support


def _rng(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[int(seed) & 0xFFFFFFFFFFFFFFFF, int(stream)]))


def lattice(nx: int, ny: int, nz: int, lo, dx: float, rng: np.random.Generator | None, jitter: float = 0.05):
    """Positions (3, nx*ny*nz), x fastest, with uniform jitter of +-jitter*dx per component."""
    ix = np.arange(nx, dtype=np.float64)
    iy = np.arange(ny, dtype=np.float64)
    iz = np.arange(nz, dtype=np.float64)
    Z, Y, X = np.meshgrid(iz, iy, ix, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    pos = (pos + 0.5) * dx
    pos += np.asarray(lo, dtype=np.float64)[:, None]
    if rng is not None and jitter > 0:
        pos += rng.uniform(-jitter * dx, jitter * dx, size=pos.shape)
    return pos


def _fill(ps: ParticleSet, pos, mass, rho, u, h, species=SPECIES_DEFAULT):
    ps.pos[:] = pos
    ps.mass[:] = mass
    ps.rho[:] = rho
    ps.u[:] = u
    ps.h[:] = h
    ps.species[:] = species


# ---------------------------------------------------------------------------------------------------
# Config 1: plain-hydro 3D Sod shock tube (the reference's CPU-runnable case)
# ---------------------------------------------------------------------------------------------------
def sod_tube(seed: int = 1, nx_left: int = 128, cross: int = 32, gamma: float = 1.4, walls: bool = True,
             kernel: int = KERNEL_WENDLAND):
    """Box [0,1]x[0,w]^2 (w = cross/(2 nx_left)), periodic in y, z. Left lattice dx = 1/(2 nx_left)
    (rho 1, P 1), right lattice 2 dx (rho 0.125, P 0.1), equal masses. Frozen Boundary slabs of
    thickness >= H beyond x = 0 and x = 1. The default (nx_left=128, cross=32) is 147,456 fluid
    particles (1/256 and 1/128 spacings) + 3,584 wall particles."""
    rng = _rng(seed)
    dxl = 0.5 / nx_left
    dxr = 2 * dxl
    w = cross * dxl
    h = h_for_spacing(dxl, kernel)
    support = 3 if kernel == KERNEL_QUINTIC else 2
    H = support * h
    nwl = int(math.ceil(H / dxl)) + 1 if walls else 0
    nwr = int(math.ceil(H / dxr)) + 1 if walls else 0
    m = 1.0 * dxl ** 3
    ul = 1.0 / ((gamma - 1) * 1.0)
    ur = 0.1 / ((gamma - 1) * 0.125)
    parts = []
    left = lattice(nx_left, cross, cross, (0, 0, 0), dxl, rng)
    right = lattice(nx_left // 2, cross // 2, cross // 2, (0.5, 0, 0), dxr, rng)
    blocks = [(left, 1.0, ul, SPECIES_DEFAULT), (right, 0.125, ur, SPECIES_DEFAULT)]
    if walls:
        wl = lattice(nwl, cross, cross, (-nwl * dxl, 0, 0), dxl, None)
        wr = lattice(nwr, cross // 2, cross // 2, (1.0, 0, 0), dxr, None)
        blocks += [(wl, 1.0, ul, SPECIES_BOUNDARY), (wr, 0.125, ur, SPECIES_BOUNDARY)]
    for pos, rho, u, sp in blocks:
        ps = ParticleSet(pos.shape[1])
        _fill(ps, pos, m, rho, u, h, sp)
        parts.append(ps)
    ps = ParticleSet.concat(parts)
    ps.id = np.arange(ps.n, dtype=np.uint64)
    lo_x = -nwl * dxl
    len_x = 1.0 + nwl * dxl + nwr * dxr
    cfg = make_config(box_lo=(lo_x, 0, 0), box_len=(len_x, w, w), periodic=(0, 1, 1), kernel=kernel, h0=h,
                      gamma=gamma, alpha_visc=1.0, beta_visc=2.0, mhd=0)
    return cfg, ps


# ---------------------------------------------------------------------------------------------------
# Config 2: Orszag-Tang vortex slab
# ---------------------------------------------------------------------------------------------------
def orszag_tang(n: int = 512, nz: int = 4, seed: int = 2, kernel: int = KERNEL_WENDLAND):
    """[0,1]^2 x [0, nz/n], fully periodic, n x n x nz lattice. rho = 25/(36 pi), P = 5/(12 pi),
    v = (-sin 2 pi y, sin 2 pi x, 0), B = B0 (-sin 2 pi y, sin 4 pi x, 0), B0 = 1/sqrt(4 pi),
    gamma = 5/3, hyperbolic cleaning on."""
    rng = _rng(seed)
    dx = 1.0 / n
    gamma = 5.0 / 3.0
    rho = 25.0 / (36.0 * math.pi)
    P = 5.0 / (12.0 * math.pi)
    B0 = 1.0 / math.sqrt(4.0 * math.pi)
    pos = lattice(n, n, nz, (0, 0, 0), dx, rng)
    ps = ParticleSet(pos.shape[1])
    h = h_for_spacing(dx, kernel)
    _fill(ps, pos, rho * dx ** 3, rho, P / ((gamma - 1) * rho), h)
    x, y = pos[0], pos[1]
    ps.vel[0] = -np.sin(2 * np.pi * y)
    ps.vel[1] = np.sin(2 * np.pi * x)
    ps.B[0] = -B0 * np.sin(2 * np.pi * y)
    ps.B[1] = B0 * np.sin(4 * np.pi * x)
    cfg = make_config(box_len=(1.0, 1.0, nz * dx), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=gamma,
                      mhd=1, cleaning=1, borve=1)
    return cfg, ps


# ---------------------------------------------------------------------------------------------------
# Config 3: MHD turbulence in a periodic cube
# ---------------------------------------------------------------------------------------------------
def _random_field(n: int, rng: np.random.Generator, kmin: float, kmax: float, slope: float):
    """Real 3D Gaussian random field on an n^3 grid with |F(k)| ~ k^slope for kmin <= |k| <= kmax."""
    k1 = np.fft.fftfreq(n) * n
    kr = np.fft.rfftfreq(n) * n
    KX, KY, KZ = np.meshgrid(k1, k1, kr, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    shell = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[shell] = kk[shell] ** slope
    re = rng.standard_normal(kk.shape)
    im = rng.standard_normal(kk.shape)
    return amp * (re + 1j * im), (KX, KY, KZ)


def mhd_turbulence(n: int = 128, seed: int = 3, mach: float = 1.0, kernel: int = KERNEL_WENDLAND):
    """Periodic unit cube, n^3 lattice, rho = 1, P = 1, gamma = 5/3. v: Gaussian random field with
    E(k) ~ k^-2 for 1 <= k <= 8 (so |v_k| ~ k^-2), rescaled to rms Mach `mach`. B = 0.1 z plus a
    divergence-free perturbation (curl of a random vector potential) of 10% rms."""
    rng = _rng(seed)
    dx = 1.0 / n
    gamma = 5.0 / 3.0
    pos = lattice(n, n, n, (0, 0, 0), dx, rng)
    ps = ParticleSet(pos.shape[1])
    h = h_for_spacing(dx, kernel)
    _fill(ps, pos, dx ** 3, 1.0, 1.0 / (gamma - 1), h)
    cs = math.sqrt(gamma)
    vel = []
    for _ in range(3):
        vk, _k = _random_field(n, rng, 1.0, 8.0, -2.0)
        vel.append(np.fft.irfftn(vk, s=(n, n, n), axes=(0, 1, 2)))
    vrms = math.sqrt(sum(float(np.mean(v ** 2)) for v in vel))
    scale = mach * cs / vrms
    Ak = []
    for _ in range(3):
        a, K = _random_field(n, rng, 1.0, 8.0, -3.0)
        Ak.append(a)
    KX, KY, KZ = K
    bk = [1j * (KY * Ak[2] - KZ * Ak[1]), 1j * (KZ * Ak[0] - KX * Ak[2]), 1j * (KX * Ak[1] - KY * Ak[0])]
    db = [np.fft.irfftn(b, s=(n, n, n), axes=(0, 1, 2)) for b in bk]
    brms = math.sqrt(sum(float(np.mean(b ** 2)) for b in db))
    bscale = 0.01 / brms
    # lattice order is x fastest: flat index = ix + n*(iy + n*iz); field arrays are [ix, iy, iz]
    for a in range(3):
        ps.vel[a] = (vel[a] * scale).transpose(2, 1, 0).ravel()
        ps.B[a] = (db[a] * bscale).transpose(2, 1, 0).ravel()
    ps.B[2] += 0.1
    cfg = make_config(box_len=(1.0, 1.0, 1.0), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=gamma, mhd=1,
                      cleaning=1, borve=1)
    return cfg, ps


# ---------------------------------------------------------------------------------------------------
# Config 4: ragged ensemble of 3D MHD Kelvin-Helmholtz runs
# ---------------------------------------------------------------------------------------------------
def kelvin_helmholtz(n_target: int, seed: int, kernel: int = KERNEL_WENDLAND):
    """Periodic box [0,1]^2 x [0, nz/n] on an n x n x nz lattice (n = round(N^(1/3))). Shear layer:
    rho 2 (by mass) for |y - 1/2| < 1/4 moving at +dV/2, rho 1 at -dV/2, dV ~ U[0.5, 1.5]; vy =
    0.01 sin(4 pi x + phase); P = 2.5; uniform Bx with plasma beta ~ U[10, 1000]; gamma = 5/3."""
    rng = _rng(seed, 1)
    n = max(8, int(round(n_target ** (1.0 / 3.0))))
    nz = max(3, int(round(n_target / (n * n))))
    dx = 1.0 / n
    gamma = 5.0 / 3.0
    dV = rng.uniform(0.5, 1.5)
    beta = rng.uniform(10.0, 1000.0)
    phase = rng.uniform(0, 2 * math.pi)
    P = 2.5
    pos = lattice(n, n, nz, (0, 0, 0), dx, rng)
    ps = ParticleSet(pos.shape[1])
    h = h_for_spacing(dx, kernel)
    band = np.abs(pos[1] - 0.5) < 0.25
    rho = np.where(band, 2.0, 1.0)
    _fill(ps, pos, rho * dx ** 3, rho, P / ((gamma - 1) * rho), h)
    ps.vel[0] = np.where(band, 0.5 * dV, -0.5 * dV)
    ps.vel[1] = 0.01 * np.sin(4 * np.pi * pos[0] + phase)
    ps.B[0] = math.sqrt(2 * P / beta)
    cfg = make_config(box_len=(1.0, 1.0, nz * dx), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=gamma, mhd=1,
                      cleaning=1, borve=1)
    return cfg, ps


def ensemble_sizes(n_runs: int = 256, seed: int = 4, lo: int = 4096, hi: int = 65536):
    rng = _rng(seed, 2)
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(n_runs)]


def kh_ensemble(n_runs: int = 256, seed: int = 4):
    return [kelvin_helmholtz(nt, seed * 100003 + r) for r, nt in enumerate(ensemble_sizes(n_runs, seed))]


# ---------------------------------------------------------------------------------------------------
# Config 5: serving mix (Brio-Wu, rotor, Orszag-Tang)
# ---------------------------------------------------------------------------------------------------
def brio_wu(n_target: int, seed: int, kernel: int = KERNEL_WENDLAND):
    """Brio & Wu shock tube in a periodic box [0,1] x [0,w]^2 (w = 16 dx): left dx, rho 1, P 1,
    By +1; right 2 dx, rho 0.125, P 0.1, By -1; Bx = 0.75; gamma = 2; equal masses. The periodic x
    boundary adds a second (mirror) interface at x = 0."""
    rng = _rng(seed, 3)
    gamma = 2.0
    nyz = 16
    nxl = max(8, int(round(n_target / (nyz * nyz * (1 + 1.0 / 16)) )))
    nxl -= nxl % 2
    dx = 0.5 / nxl
    w = nyz * dx
    left = lattice(nxl, nyz, nyz, (0, 0, 0), dx, rng)
    right = lattice(nxl // 2, nyz // 2, nyz // 2, (0.5, 0, 0), 2 * dx, rng)
    h = h_for_spacing(dx, kernel)
    m = dx ** 3
    a = ParticleSet(left.shape[1])
    _fill(a, left, m, 1.0, 1.0 / ((gamma - 1) * 1.0), h)
    a.B[1] = 1.0
    b = ParticleSet(right.shape[1])
    _fill(b, right, m, 0.125, 0.1 / ((gamma - 1) * 0.125), h)
    b.B[1] = -1.0
    ps = ParticleSet.concat([a, b])
    ps.id = np.arange(ps.n, dtype=np.uint64)
    ps.B[0] = 0.75
    cfg = make_config(box_len=(1.0, w, w), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=gamma, mhd=1,
                      cleaning=1, borve=1)
    return cfg, ps


def rotor(n_target: int, seed: int, kernel: int = KERNEL_WENDLAND):
    """Magnetised rotor (Balsara & Spicer 1999) in a periodic slab [0,1]^2 x [0, 4 dx]: disc r < 0.1
    with rho 10 (by mass) rotating rigidly with edge speed 2, ambient rho 1 at rest, P = 1,
    Bx = 5/sqrt(4 pi), gamma = 1.4."""
    rng = _rng(seed, 4)
    gamma = 1.4
    n = max(16, int(round(math.sqrt(n_target / 4))))
    dx = 1.0 / n
    pos = lattice(n, n, 4, (0, 0, 0), dx, rng)
    rx, ry = pos[0] - 0.5, pos[1] - 0.5
    r = np.sqrt(rx ** 2 + ry ** 2)
    disc = r < 0.1
    rho = np.where(disc, 10.0, 1.0)
    ps = ParticleSet(pos.shape[1])
    h = h_for_spacing(dx, kernel)
    _fill(ps, pos, rho * dx ** 3, rho, 1.0 / ((gamma - 1) * rho), h)
    omega = 2.0 / 0.1
    ps.vel[0] = np.where(disc, -omega * ry, 0.0)
    ps.vel[1] = np.where(disc, omega * rx, 0.0)
    ps.B[0] = 5.0 / math.sqrt(4 * math.pi)
    cfg = make_config(box_len=(1.0, 1.0, 4 * dx), periodic=(1, 1, 1), kernel=kernel, h0=h, gamma=gamma, mhd=1,
                      cleaning=1, borve=1)
    return cfg, ps


def orszag_tang_n(n_target: int, seed: int, kernel: int = KERNEL_WENDLAND):
    n = max(16, int(round(math.sqrt(n_target / 4))))
    return orszag_tang(n=n, nz=4, seed=seed, kernel=kernel)


SERVING_PROBLEMS = {"brio_wu": brio_wu, "rotor": rotor, "orszag_tang": orszag_tang_n}
