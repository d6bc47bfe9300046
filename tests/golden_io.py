"""Loading the golden fixtures of tests/golden/ (test infrastructure)."""
from __future__ import annotations

import os

import numpy as np

from golden.make_golden import CASES, config_from_json, particles_from

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASE_NAMES = sorted(CASES)


def load_case(name: str):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    return config_from_json(z["config"]), particles_from(z), z


def load_kernels():
    return np.load(os.path.join(GOLDEN, "kernels_closed_form.npz"), allow_pickle=False)
