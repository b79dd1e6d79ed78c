"""Grid excluded volume with stretchable dummies (HistogramManagerMTFFGridScalableExv,
HistogramManagerMTFFGridScalableExv.cpp:16-244; CompositeDistanceHistogramFFGridScalableExv.cpp:9-19).

At stretch k the dummies' float coordinates are multiplied by k and their form
factor is ExvFormFactor(k^3). Pins: k = 1 is the plain FFGrid structure; the
HIP histograms at k match the oracle's FFGrid histograms of the stretched
dummies bit for bit; and the Debye sum fixes I_xx(q; k) = k^6 I_xx(kq; 1) for
the dummies alone (stretching every distance by k and the form factor's width
by 1/k), which holds on the weighted bin centres to the binning error.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def structure(n=600, m=150, seed=41):
    r = 12.0 * (n / 600) ** (1 / 3)
    g = A.synth_globule(seed, n, (r, r, r), 2.0)
    w = A.synth_shell(seed + 1, m, (r, r, r), 3.0)[:, :3].astype(np.float64) if m else None
    return g[:, :3].astype(np.float64), g[:, 3].view(np.int32).copy(), w


def test_unit_stretch_is_the_grid_structure():
    xyz, ff, wxyz = structure()
    sc = A.ScalableGridStructure(xyz, ff, wxyz)
    h1 = sc.calculate(1.0)
    h0 = A.GridStructure(xyz, ff, wxyz).calculate()
    for w in (A.AA, A.AW, A.WW, A.PTOT):
        for x, y in zip(h1.get(w), h0.get(w)):
            assert np.array_equal(x, y)
    q = np.linspace(0, 0.5, 60)
    assert np.array_equal(h1.debye_transform(q), h0.debye_transform(q))


@pytest.mark.parametrize("k", [0.95, 1.06])
def test_stretched_dummies_match_oracle(k):
    xyz, ff, wxyz = structure()
    h = A.ScalableGridStructure(xyz, ff, wxyz).calculate(k)
    ref = O.hist_ffgrid(xyz, ff, wxyz, nthreads=4, exv_scale=k)
    v, c, _ = h.get(A.AA)
    assert np.array_equal(c, ref["aa"]["count"])
    v, c, _ = h.get(A.AW)
    assert np.array_equal(c, ref["aw"]["count"])
    q = np.linspace(0.0, 0.5, 101)
    I = h.debye_transform(q)
    I_ref, _ = O.debye_ffgrid(ref, q, 0.25, grid_exv_volume=k ** 3)
    assert np.max(np.abs(I - I_ref) / np.abs(I_ref)) <= 1e-4


def test_dummy_profile_scaling_law():
    xyz, ff, wxyz = structure(900, 0)
    sc = A.ScalableGridStructure(xyz, ff, None)
    q = np.linspace(0.02, 0.4, 40)
    for k in (0.92, 1.08):
        _, p_k = sc.calculate(k).debye_transform(q, profiles=True)
        _, p_1 = sc.calculate(1.0).debye_transform(k * q, profiles=True)
        xx_k, xx_1 = p_k[3], p_1[3]
        # relative to the profile's scale (it passes through a zero near q = 0.33)
        assert np.max(np.abs(xx_k - k ** 6 * xx_1)) < 1e-4 * np.max(np.abs(k ** 6 * xx_1))


def test_fit_excluded_volume_stretch():
    xyz, ff, wxyz = structure(500, 120, seed=43)
    sc = A.ScalableGridStructure(xyz, ff, wxyz)
    q = np.linspace(0.02, 0.45, 60)
    I = 3e-3 * sc.calculate(1.04).debye_transform(q) + 0.2
    f = A.Fitter(sc, q, I, 0.01 * I, enabled=A.FIT_EXCLUDED_VOLUME)
    r = f.fit(A.MINI_GOLDEN)
    assert abs(r["params"][0] - 1.04) < 2e-3
    assert r["chi2"] < 0.1  # 60 points with 1 % errors: residuals far below the errors
