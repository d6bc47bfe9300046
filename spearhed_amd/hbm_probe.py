"""Run the known-size device copy (sph_probe_copy) for HBM counter calibration.

Usage: python -m spearhed_amd.hbm_probe [MiB] [reps]. Under rocprofv3 --pmc FETCH_SIZE (or
WRITE_SIZE) the per-dispatch reading of k_probe_copy divided by the copied bytes is the counter's
scale factor; a copy of B bytes reads B and writes B.
"""
from __future__ import annotations

import ctypes as C
import json
import sys

try:
    from . import capi
except ImportError:  # run as a script
    import os as _os
    sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
    from spearhed_amd import capi


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    lib = capi.load()
    ms = C.c_double()
    nbytes = mib * (1 << 20)
    capi.check(lib.sph_probe_copy(nbytes, reps, C.byref(ms)), None, "sph_probe_copy")
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    print(json.dumps({"bytes_copied": nbytes, "ms_per_copy": ms.value, "read_plus_write_GBps": gbs}))


if __name__ == "__main__":
    main()
