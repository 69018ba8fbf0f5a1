import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def golden_manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["cases"]


def forced_manifest():
    """Golden cases of the reference's StratosphereRK4 (forced step)."""
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["forced_cases"]


def voronoi_manifest():
    """Golden cases of the reference's Voronoi BVERK4."""
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["voronoi_cases"]


def voronoi_step_inputs(g, c, s):
    """Inputs of step s (1-based) of a Voronoi golden case: the state after
    step s-1, the areas the sum weighs (the initial tessellation's with
    UseSavedArea, else the current ones) and the current areas."""
    p = f"s{s - 1}_"
    sum_area = g["s0_area"] if c["saved_area"] else g[p + "area"]
    return g[p + "x"], g[p + "y"], g[p + "z"], g[p + "relvort"], sum_area, g[p + "area"]


def load_golden(name):
    """Arrays of one golden case (written by tests/golden/make_golden.py)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def normwise_rel(a, b):
    """max_i |a_i - b_i| / max_i |b_i| over points (rows) -- the parity metric."""
    a = np.asarray(a, dtype=np.float64).reshape(len(b), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    den = np.max(np.linalg.norm(b, axis=1))
    return float(np.max(np.linalg.norm(a - b, axis=1)) / den) if den > 0 else float(np.max(np.abs(a - b)))


@pytest.fixture(scope="session")
def manifest():
    return golden_manifest()
