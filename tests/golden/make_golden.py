"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py [case ...]).

Two kinds of fixture:
  * kernels_closed_form.npz -- W(r, h) and dW/dr(r, h) of the cubic M4, quintic M5 and Wendland C2
    kernels, evaluated in float64 numpy from the published closed forms (Monaghan & Lattanzio 1985;
    Morris 1996; Wendland 1995 with the support-2h scaling of Dehnen & Aly 2012), independently of
    the oracle. The reference's shape functions: sph/CubicSplineKernel.hpp:47-137,
    sph/QuinticSplineKernel.hpp:46-160.
  * <case>.npz -- seeded inputs of a small run (<= 4k particles) and the FP64 oracle's outputs:
    supercell keys and starts, neighbour counts at each precision, density, forces and the FP32
    summation scales after one resort/density/forces pass; then the conserved sums and dt after
    prime + 3 full steps, and the final state keyed by particle id.

Inputs and outputs are stored in the dtypes the C ABI uses; the run configuration is stored as
JSON text. Everything loads with numpy.load(allow_pickle=False).
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_ref as O  # noqa: E402
from spearhed_amd import capi, ic  # noqa: E402
from spearhed_amd.particles import ParticleSet  # noqa: E402

STEPS = 3
INPUT_FIELDS = ("id", "mask", "species", "pos", "mass", "vel", "rho", "h", "u", "B", "psi")


def config_to_json(cfg) -> str:
    d = {}
    for name, _ in cfg._fields_:
        v = getattr(cfg, name)
        d[name] = list(v) if hasattr(v, "__len__") else v
    return json.dumps(d)


def config_from_json(text: str):
    d = json.loads(str(text))
    c = capi.RunConfig()
    for name, _ in c._fields_:
        v = d[name]
        if isinstance(v, list):
            getattr(c, name)[:] = v
        else:
            setattr(c, name, v)
    return c


def particles_from(z) -> ParticleSet:
    ps = ParticleSet(int(z["id"].shape[0]))
    for f in INPUT_FIELDS:
        setattr(ps, f, z[f].copy())
    return ps


# -- closed-form kernels -------------------------------------------------------------------------
def closed_form(kernel: int, r: np.ndarray, h: float):
    q = r / h
    if kernel == capi.KERNEL_CUBIC:
        s = 1.0 / math.pi
        f = np.where(q < 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))
        g = np.where(q < 1, 2.25 * q ** 2 - 3 * q, np.where(q < 2, -0.75 * (2 - q) ** 2, 0.0))
    elif kernel == capi.KERNEL_QUINTIC:
        s = 1.0 / (120.0 * math.pi)
        a, b, c = np.clip(3 - q, 0, None), np.clip(2 - q, 0, None), np.clip(1 - q, 0, None)
        f = a ** 5 - 6 * b ** 5 + 15 * c ** 5
        g = -5 * a ** 4 + 30 * b ** 4 - 75 * c ** 4
    else:
        s = 21.0 / (16.0 * math.pi)
        t = np.clip(1 - 0.5 * q, 0, None)
        f = t ** 4 * (1 + 2 * q)
        g = -5 * q * t ** 3
    return s / h ** 3 * f, s / h ** 4 * g


def make_kernels(path):
    out = {}
    for kernel in (capi.KERNEL_CUBIC, capi.KERNEL_QUINTIC, capi.KERNEL_WENDLAND):
        sup = 3.0 if kernel == capi.KERNEL_QUINTIC else 2.0
        for hi, h in enumerate((1.0, 0.5, 0.0371)):
            q = np.concatenate([np.linspace(0, sup, 97), [0.999, 1.0, 1.001, 1.999, 2.0, 2.001, sup - 1e-3]])
            r = q * h
            w, dw = closed_form(kernel, r, h)
            out[f"k{kernel}_h{hi}_r"] = r
            out[f"k{kernel}_h{hi}_w"] = w
            out[f"k{kernel}_h{hi}_dwdr"] = dw
            out[f"k{kernel}_h{hi}_h"] = np.array(h)
    np.savez_compressed(path, **out)


# -- oracle runs ---------------------------------------------------------------------------------
CASES = {
    # name: (builder, precision used for the stepped part)
    "turbulence_12": lambda: ic.mhd_turbulence(n=12, seed=101),
    "sod_walls_small": lambda: ic.sod_tube(seed=102, nx_left=24, cross=4),
    "kelvin_helmholtz_3k": lambda: ic.kelvin_helmholtz(3000, 103),
    "orszag_tang_2k": lambda: ic.orszag_tang_n(2300, 104),
    "cubic_turbulence_10": lambda: ic.mhd_turbulence(n=10, seed=105, kernel=capi.KERNEL_CUBIC),
    "quintic_turbulence_9": lambda: ic.mhd_turbulence(n=9, seed=106, kernel=capi.KERNEL_QUINTIC),
    # the serving mix's two other problem types (gamma = 2 and a discontinuous B_y; a dense rotating disc)
    "brio_wu_3k": lambda: ic.brio_wu(3000, 107),
    "rotor_3k": lambda: ic.rotor(3000, 108),
}


def make_case(name, path):
    cfg, ps = CASES[name]()
    assert ps.n <= 4096, (name, ps.n)
    out = {"config": np.array(config_to_json(cfg))}
    for f in INPUT_FIELDS:
        out[f] = getattr(ps, f)
    # one pass of the stages, all precisions (neighbour counts depend on the precision's r^2)
    for prec in ("plain", "mixed", "f64"):
        o = O.OracleRun(cfg, ps, prec, scales=(prec == "f64"))
        cs = o.resort()
        o.density()
        o.forces()
        order = o.by_id()
        out[f"{prec}_nnb"] = o.nnb[order].copy()
        if prec == "f64":
            out["f64_cell_start"] = cs.copy()
            out["f64_key_by_id"] = O.bin_keys(cfg, "f64", ps.pos)
            for f in ("rho", "dudt", "dpsidt", "divb", "vsig"):
                out[f"f64_{f}"] = getattr(o, f)[order].copy()
            for f in ("dvdt", "dbdt"):
                out[f"f64_{f}"] = getattr(o, f)[:, order].copy()
            for f in ("sc_rho", "sc_a", "sc_du", "sc_db", "sc_dpsi", "sc_divb"):
                out[f"f64_{f}"] = getattr(o, f)[order].copy()
        if prec in ("plain", "mixed"):
            out[f"{prec}_key_by_id"] = O.bin_keys(cfg, prec, ps.pos)
    # prime + steps with the FP64 oracle
    o = O.OracleRun(cfg, ps, "f64")
    c0 = o.prime()
    cons = [c0]
    for _ in range(STEPS):
        cons.append(o.step(1))
    keys = ("mass", "energy", "dt_next", "time", "c_h")
    for k in keys:
        out[f"steps_{k}"] = np.array([c[k] for c in cons], dtype=np.float64)
    out["steps_momentum"] = np.array([c["momentum"] for c in cons], dtype=np.float64)
    order = o.by_id()
    out["steps_pos"] = o.pos[:, order].copy()
    for f in ("vel", "B"):
        out[f"steps_{f}"] = getattr(o, f)[:, order].copy()
    for f in ("u", "psi", "rho"):
        out[f"steps_{f}"] = getattr(o, f)[order].copy()
    np.savez_compressed(path, **out)
    return ps.n


def main(names=None):
    if not names:
        make_kernels(os.path.join(HERE, "kernels_closed_form.npz"))
    for name in names or CASES:
        n = make_case(name, os.path.join(HERE, f"{name}.npz"))
        print(f"{name}: {n} particles")


if __name__ == "__main__":
    main(sys.argv[1:])  # case names (default: every fixture)
