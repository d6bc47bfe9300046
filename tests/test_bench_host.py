"""bench.py's host-side contract (CPU only): the one-GPU guard, the two CPU baselines and the
measured-traffic lookup. The GPU legs run on an MI355X (`python bench.py`)."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

from spearhed_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr and res.stdout == ""


def test_cpu_baselines_report_the_documented_fields(oracle_lib):
    cfg, ps = ic.mhd_turbulence(n=12, seed=9)
    ref = bench.cpu_reference_structured(cfg, ps, 0.2, 2)
    binned = bench.cpu_baseline(cfg, ps, 0.2)
    for rec in (ref, binned):
        assert rec["value"] > 0 and rec["unit"] == "particle-steps/s" and rec["kind"] == "port"
        assert rec["cores"] >= 1 and f"{ps.n:,}" in rec["sample"]
    assert "region traversal" in ref["sample"] and "64-slot frames" in ref["sample"]
    assert "27-supercell" in binned["sample"]


def test_fp64_error_report(oracle_lib):
    """The FP64 comparison on a stand-in for the GPU's first pass (the mixed FP32 oracle, whose counts
    the GPU matches bit for bit): errors within the FP32 bounds, mismatching counts counted."""
    import oracle_ref as O
    cfg, ps = ic.mhd_turbulence(n=12, seed=9)
    o = O.OracleRun(cfg, ps, "mixed")
    o.resort()
    o.density()
    o.forces()
    g = o.to_particles().sorted_by_id()
    rep = bench.fp64_error(cfg, ps, g)
    for key in ("rho", "dvdt", "dudt", "dbdt", "divb"):
        assert 0.0 <= rep[key] < 1e-5, (key, rep[key])
    assert rep["neighbour_count_mismatches_fp64"] >= 0
    assert rep["neighbour_counts_equal_fp64"] == (rep["neighbour_count_mismatches_fp64"] == 0)


def test_measured_traffic_only_for_the_profiled_workload():
    rec = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))["k_forces"]
    assert bench.measured_traffic("k_forces", rec["particles"], rec["tuning"]) == rec["bytes_per_launch"]
    other = dict(rec["tuning"], force_chunk=128 if rec["tuning"]["force_chunk"] != 128 else 64)
    assert bench.measured_traffic("k_forces", rec["particles"], other) is None
    assert bench.measured_traffic("k_forces", rec["particles"] + 1, rec["tuning"]) is None
    assert bench.measured_traffic("k_density", rec["particles"], rec["tuning"]) is None


def test_step_roofline_sums_stage_bounds():
    """At 2M particles and 60 pairs each: the force pass is FP32-bound (13.6 GFLOP at 157.3 TFLOP/s,
    0.087 ms, against 0.033 ms for its bytes), the HBM-bound stages cost bytes / 8 TB/s."""
    n, pairs = 2_097_152, 60 * 2_097_152
    r = bench.step_roofline(n, pairs, 1.85)
    st = r["stage_sol_ms"]
    assert st["forces"] == pytest.approx((108 * pairs + 80 * n) / 157.3e12 * 1e3)
    assert st["drift_bin"] == pytest.approx(70 * n / 8e12 * 1e3)
    assert r["sol_ms"] == pytest.approx(sum(st.values()))
    assert r["frac"] == pytest.approx(r["sol_ms"] / 1.85) and 0 < r["frac"] < 1
