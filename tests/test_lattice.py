"""Dummy-dummy pair histograms by lattice correlation.

The grid's dummies sit on integer coordinates, so their body pairs can be
counted from the FFT autocorrelation of the occupancy grid (cross-correlation
for interior x surface) instead of pair by pair; the reference's scalar-rule
tail pairs keep their per-pair evaluation. Every count and fixed-point offset
sum is an integer in both paths, so the two paths must give identical
histograms: these tests compare them bit for bit (lattice = 2 forces the
lattice path, 0 turns it off), and check that the automatic choice takes the
lattice path where it pays. The histograms themselves are held to the oracle
by the FFGrid and FFGridSurface parity tests, which run with the default
(automatic) setting.
"""
import numpy as np
import pytest

import ausaxs_amd as A

pytestmark = pytest.mark.gpu

GRID_DISTS = (A.AA, A.AW, A.WW, A.PTOT, A.PTOT_AX, A.XX)
SURFACE_DISTS = GRID_DISTS + (A.XX_INTERIOR, A.XX_SURFACE, A.XX_CROSS, A.WX_INTERIOR, A.WX_SURFACE,
                              A.AX_INTERIOR, A.AX_SURFACE)


def structure(seed, n, m):
    r = 20.0 * (n / 2000.0) ** (1 / 3)
    g = A.synth_globule(seed, n, (r, r, r), 2.0)
    w = A.synth_shell(seed + 1, m, (r, r, r), 3.0)[:, :3].astype(np.float64)
    return g[:, :3].astype(np.float64), g[:, 3].view(np.int32).copy(), w


def run(cls, xyz, ff, w, lattice, **kw):
    st = cls(xyz, ff, w, settings=A.default_settings(lattice=lattice, **kw))
    st.set_timing(True)
    h = st.calculate()
    return h, st.last_timing()


def same(h1, h2, dists):
    assert h1.max_bin == h2.max_bin
    for which in dists:
        for x, y in zip(h1.get(which), h2.get(which)):
            assert np.array_equal(x, y), which


@pytest.mark.parametrize("seed,n,m", [(1, 300, 80), (2, 1200, 300)])
def test_grid_lattice_equals_pair_kernel(seed, n, m):
    xyz, ff, w = structure(seed, n, m)
    h2, t2 = run(A.GridStructure, xyz, ff, w, 2)
    h0, t0 = run(A.GridStructure, xyz, ff, w, 0)
    assert t2["lattice_pairs"] > 0 and t0["lattice_pairs"] == 0
    assert t2["pairs"] == t0["pairs"] and t2["kernel_pairs"] < t0["kernel_pairs"]
    same(h2, h0, GRID_DISTS)
    q = np.linspace(0, 0.5, 50)
    assert np.array_equal(h2.debye_transform(q), h0.debye_transform(q))


def test_surface_lattice_equals_pair_kernel():
    xyz, ff, w = structure(3, 900, 250)
    h2, t2 = run(A.GridSurfaceStructure, xyz, ff, w, 2)
    h0, t0 = run(A.GridSurfaceStructure, xyz, ff, w, 0)
    assert t2["lattice_pairs"] > 0 and t0["lattice_pairs"] == 0
    same(h2, h0, SURFACE_DISTS)


@pytest.mark.parametrize("bin_width", [0.25, 0.1])
def test_automatic_choice_on_a_larger_grid(bin_width):
    # 6,000 atoms: ~1.4e5 dummies, 1e10 dummy pairs against a 128^3 transform
    xyz, ff, w = structure(4, 6000, 1500)
    h1, t1 = run(A.GridStructure, xyz, ff, w, 1, bin_width=bin_width)
    h0, t0 = run(A.GridStructure, xyz, ff, w, 0, bin_width=bin_width)
    print(f"bin width {bin_width}: lattice {t1['total_ms']:.2f} ms ({t1['lattice_pairs']:.3g} pairs by correlation), "
          f"pair kernel {t0['total_ms']:.2f} ms")
    assert t1["lattice_pairs"] > 0.8 * t1["pairs"]
    same(h1, h0, GRID_DISTS)


def test_small_sets_stay_on_the_pair_kernel():
    # below the automatic threshold nothing changes; forced, at most three
    # dummies leave no body pairs for the lattice path
    xyz, ff, w = structure(5, 200, 40)
    _, t1 = run(A.GridStructure, xyz, ff, w, 1)
    assert t1["lattice_pairs"] == 0
    one = np.zeros((1, 3))
    h2, t2 = run(A.GridStructure, one, np.array([3], np.int32), None, 2)
    assert t2["lattice_pairs"] > 0  # 33 dummies around one carbon
    h0, _ = run(A.GridStructure, one, np.array([3], np.int32), None, 0)
    same(h2, h0, GRID_DISTS)


def test_lattice_setting_is_checked():
    xyz, ff, w = structure(5, 50, 10)
    with pytest.raises(A.AusaxsError):
        A.GridStructure(xyz, ff, w, settings=A.default_settings(lattice=3))


@pytest.mark.parametrize("seed", range(8))
def test_lattice_path_on_small_random_structures(seed):
    # small, ragged cases: dummy sets whose sizes leave every tail length, few
    # surface dummies (short cross columns), both bin widths
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 120))
    m = int(rng.integers(0, 40))
    xyz, ff, w = structure(200 + seed, max(n, 2), max(m, 1))
    w = w[:m]
    bw = (0.25, 0.1)[seed % 2]
    for cls, dists in ((A.GridStructure, GRID_DISTS), (A.GridSurfaceStructure, SURFACE_DISTS)):
        h2, t2 = run(cls, xyz, ff, w, 2, bin_width=bw)
        h0, _ = run(cls, xyz, ff, w, 0, bin_width=bw)
        same(h2, h0, dists)


@pytest.mark.parametrize("n", [40, 150, 400, 900, 1800])
def test_lattice_transform_lengths(n):
    # growing globules: occupied boxes whose padded lengths (the shortest
    # 2^a 3^b 5^c >= 2E - 1 per axis) take every radix mix of the transform,
    # odd lengths included; one set (autocorrelation) and interior x surface
    # (cross-correlation)
    xyz, ff, w = structure(300 + n, n, max(1, n // 4))
    for cls, dists in ((A.GridStructure, GRID_DISTS), (A.GridSurfaceStructure, SURFACE_DISTS)):
        h2, t2 = run(cls, xyz, ff, w, 2)
        h0, _ = run(cls, xyz, ff, w, 0)
        assert t2["lattice_pairs"] > 0
        same(h2, h0, dists)


@pytest.mark.parametrize("sep", [90.0, 240.0, 430.0])
def test_lattice_long_lines(sep):
    # two small globules far apart on x (and a third off axis): occupied boxes
    # hundreds of lattice steps long on one axis, so the x lines run to
    # lengths of several hundred (few lines per workgroup, long radix-3 and
    # radix-5 stage chains) while the other axes stay short
    xyz, ff, w = structure(500 + int(sep), 150, 40)
    shift = np.array([[sep, 0.0, 0.0]])
    third = np.array([[sep / 3, 25.0, -10.0]])
    xyz = np.concatenate([xyz, xyz + shift, xyz[:50] + third])
    ff = np.concatenate([ff, ff, ff[:50]])
    w = np.concatenate([w, w + shift])
    h2, t2 = run(A.GridStructure, xyz, ff, w, 2)
    h0, _ = run(A.GridStructure, xyz, ff, w, 0)
    assert t2["lattice_pairs"] > 0
    same(h2, h0, GRID_DISTS)
