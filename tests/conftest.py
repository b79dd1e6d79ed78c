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


def delta_exp(bin_width):
    """Exponent E of the device's offset unit 2^-E A: the largest E with
    2^E / inv_w <= 2^20, inv_w = float(1 / bin_width)."""
    return int(np.frexp(np.float32(1.0 / bin_width))[1]) + 19


def axis_ok(d, d_ref, bin_width, count=None):
    """Bin centres from the device against the oracle's. Each pair's offset on
    the device is that of the real root in units of 2^-E A (E = delta_exp:
    2^-23 A at 0.1 A bins, 2^-22 at 0.25), within 1.07 units; the oracle sums
    correctly rounded float roots, within half an ulp (2^-24 relative) of the
    real ones. A bin of n pairs therefore agrees to 1.07 units plus 2^-24 |d|,
    and the random part of that half ulp shrinks as 1 / sqrt(n): bins with
    many pairs are held to 8 / sqrt(n) of it, but not below a quarter. The
    floor is the rounding's own bias: where the float grid of d2 steps
    through the roots at a simple ratio of their ulp (d near 64, 128, 160,
    192 A ...), the correctly rounded roots of a bin are not spread evenly
    around the real ones, and their mean was measured 0.1 of half an ulp off
    on a 250k-atom structure."""
    d, d_ref = np.asarray(d, np.float64), np.asarray(d_ref, np.float64)
    unit = 1.07 * 2.0 ** -delta_exp(bin_width)
    half_ulp = 2.0 ** -24 * np.abs(d_ref)
    if count is not None:
        n = np.maximum(np.asarray(count, np.float64), 1.0)
        half_ulp = half_ulp * np.clip(8.0 / np.sqrt(n), 0.25, 1.0)
    return np.all(np.abs(d - d_ref) <= unit + half_ulp + 1e-12 * np.abs(d_ref))
