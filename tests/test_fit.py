"""Device FP64 least squares and the exact Debye sum against the oracle."""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def test_lsq_batch_matches_oracle():
    rng = np.random.default_rng(3)
    n, m = 300, 16
    data_x = rng.uniform(1, 1e4, n)
    err = rng.uniform(1, 20, n)
    y = 0.3 * data_x + 12 + rng.normal(0, 1, n) * err
    models = data_x[None, :] * rng.uniform(0.5, 2.0, (m, 1)) + rng.normal(0, 5, (m, n))
    out = A.lsq(models, y, err)
    for k in range(m):
        ref = O.lsq(models[k], y, err)
        np.testing.assert_allclose(out[k], ref, rtol=1e-10)
    one = A.lsq(models[0], y)
    np.testing.assert_allclose(one, O.lsq(models[0], y), rtol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 7, 301])
def test_exact_debye_matches_oracle(n):
    rng = np.random.default_rng(n)
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.normal(size=(n, 3)) * 8
    p[:, 3] = rng.integers(1, 9, n)
    q = np.linspace(0.0, 0.8, 40)
    I = A.exact_debye(p, q)
    I_ref = O.exact_debye(p, q)
    np.testing.assert_allclose(I, I_ref, rtol=1e-11)


def test_exact_debye_agrees_with_binned():
    # the binned Debye of the plain manager (0.01 A weighted bins, FP64
    # contraction) approaches the exact sum; the residual is the binning error
    # (sinc at the bin's mean distance instead of the mean of the sincs)
    atoms = A.synth_globule(2, 800, (14, 14, 14), 2.0)
    xyzw = atoms.copy()
    xyzw[:, 3] = 1.0
    q = np.linspace(0.01, 0.5, 30)
    I_exact = A.exact_debye(xyzw, q)
    s = A.default_settings(bin_width=0.01, bin_count=8000, debye_terms=0)
    I_bin = A.calculate_all(A.MODE_SIMPLE, xyzw, None, settings=s).debye_transform(q)
    err = np.max(np.abs(I_bin - I_exact) / np.abs(I_exact))
    err0 = np.max(np.abs(I_bin - I_exact)) / I_exact[0]
    print(f"binned vs exact: max rel err {err:.2e}, max err / I(0) {err0:.2e}")
    # binning error: ~1e-8 of I(0), 1e-4 relative where I(q) has fallen 1000x
    assert err0 < 1e-7 and err < 1e-3
