"""GPU parity at the benchmark configurations' full sizes.

The oracle's binned C restatement (OpenMP over supercells) finishes one density and force pass over
2M particles in about a second, so the full-size cases are compared with it directly, with the
same contracts as the small parity tests:
  * keys, cell starts and the (cell, id) permutation bit-exact (2,097,152-particle MHD turbulence);
  * per-particle neighbour counts bit-exact, density <= 1e-5 relative, derivatives <= 1e-5 of their
    summation scale: MHD turbulence 128^3 (2,097,152), Orszag-Tang 512 x 512 x 4 (1,048,576), the Sod
    tube at the stated spacings with its walls (151,040);
  * the 256-run Kelvin-Helmholtz ensemble (about 6M particles in one context, one launch per stage):
    a sample of its runs, each against the oracle run alone.
"""
from __future__ import annotations

import functools

import numpy as np
import pytest

import oracle_ref
from parity import compare_forces, gpu_stages, oracle_stages
from spearhed_amd import ic
from spearhed_amd.stepper import Context

pytestmark = pytest.mark.gpu

GPREC = {"plain": "fp32", "mixed": "mixed"}


@functools.lru_cache(maxsize=None)
def _turbulence_2m():
    return ic.mhd_turbulence(n=128, seed=3)


def test_turbulence_2m_sort_bit_exact():
    cfg, ps = _turbulence_2m()
    print(f"full-size sort n={ps.n}", flush=True)
    with Context([cfg], [ps.n], precision="mixed") as ctx:
        ctx.upload(0, ps)
        ctx.resort()
        key, cs = ctx.sort_info(0)
        g = ctx.download(0)
    okey = oracle_ref.bin_keys(cfg, "mixed", ps.pos)
    perm, ocs = oracle_ref.sort_perm(okey, ps.id, oracle_ref.grid(cfg).ncell)
    assert np.array_equal(cs, ocs), "cell starts differ"
    assert np.array_equal(g.id, ps.id[perm]), "sort permutation differs"
    assert np.array_equal(key, okey[perm]), "supercell keys differ"


@pytest.mark.parametrize("prec", ["mixed", "plain"])
def test_turbulence_2m_forces(prec):
    cfg, ps = _turbulence_2m()
    print(f"full-size turbulence n={ps.n} prec={prec}", flush=True)
    g = gpu_stages(cfg, ps, GPREC[prec])
    o = oracle_stages(cfg, ps, prec)
    compare_forces(g, o, mhd=True)


def test_orszag_tang_1m_forces():
    cfg, ps = ic.orszag_tang(n=512, nz=4, seed=2)
    print(f"full-size Orszag-Tang n={ps.n}", flush=True)
    g = gpu_stages(cfg, ps, "mixed")
    o = oracle_stages(cfg, ps, "mixed")
    compare_forces(g, o, mhd=True)


@pytest.mark.parametrize("prec", ["mixed", "plain"])
def test_sod_full_forces(prec):
    cfg, ps = ic.sod_tube(seed=1)
    print(f"full-size Sod n={ps.n} prec={prec}", flush=True)
    g = gpu_stages(cfg, ps, GPREC[prec])
    o = oracle_stages(cfg, ps, prec)
    compare_forces(g, o, mhd=False)


def test_ensemble_256_sampled_runs_against_the_oracle():
    runs = ic.kh_ensemble(256, seed=4)
    sizes = [p.n for _, p in runs]
    print(f"full-size ensemble runs={len(runs)} particles={sum(sizes)}", flush=True)
    sample = [0, 1, int(np.argmin(sizes)), int(np.argmax(sizes)), 97, 128, 200, 255]
    with Context([c for c, _ in runs], sizes, precision="mixed") as ctx:
        for r, (_, ps) in enumerate(runs):
            ctx.upload(r, ps)
        ctx.resort()
        ctx.density()
        ctx.forces()
        ctx.sync()
        got = {r: ctx.download(r).sorted_by_id() for r in sample}
    for r in sample:
        cfg, ps = runs[r]
        o = oracle_stages(cfg, ps, "mixed")
        compare_forces(got[r], o, mhd=True)
