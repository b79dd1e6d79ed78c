import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")


@pytest.fixture(scope="session")
def golden_carbon():
    d = np.load(ROOT / "tests" / "golden" / "carbon_sphere.npz")
    return d["xyz"], d["q"], d["I"]


def compare_hist_approx(p1, p2, abs_tol=1e-6, rel=1e-3):
    """hist_test_helper.h:49-74: elementwise approx, else a +-1 index window sum."""

    def approx(a, b):
        return abs(a - b) <= max(abs_tol, rel * max(abs(a), abs(b)))

    n = min(len(p1), len(p2))
    for i in range(n):
        if not approx(p1[i], p2[i]):
            lo, hi = max(0, i - 1), min(n, i + 2)
            if not approx(sum(p1[lo:hi]), sum(p2[lo:hi])):
                return False, i
    return True, -1


def cube_atoms(ff=3):
    """SimpleCube (hist_test_helper.h:87-162): 8 points at (+-1, +-1, +-1)."""
    pts = np.array([[x, y, z] for z in (-1, 1) for x in (-1, 1) for y in (-1, 1)], dtype=np.float32)
    # reference order: (-1,-1,-1), (-1,1,-1), (1,-1,-1), (1,1,-1), then z = +1
    return pts


def globule(n, radius, seed, r_excl=2.0):
    import ausaxs_amd as A

    return A.synth_globule(seed, n, (radius, radius, radius), r_excl)
