"""Fitting a histogram to measured I(q) (SmartFitter, SmartFitter.cpp:1-252).

CPU tests pin the oracle's pieces: math::CubicSpline against SciPy's natural
cubic spline (an independent implementation of the same end conditions) and
mini::Golden against a parabola with a known minimum and the evaluation count
its loop fixes. GPU tests compare the batched chi2 evaluation with the oracle
(FP64 Debye, spline, linear least squares) and check that noise-free data
made from the model give back the parameters they were made with.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

F32 = np.float32


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_spline_is_the_natural_cubic_spline(impl):
    from scipy.interpolate import CubicSpline
    f = O.spline if impl == "oracle" else A.cubic_spline
    rng = np.random.default_rng(1)
    for n in (4, 7, 40, 200):
        x = np.sort(rng.uniform(0, 5, n))
        y = np.sin(2 * x) + rng.normal(0, 0.1, n)
        z = np.linspace(x[0] - 0.3, x[-1] + 0.3, 97)  # both extrapolation ends
        ref = CubicSpline(x, y, bc_type="natural")(z)
        np.testing.assert_allclose(f(x, y, z), ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(f(x, y, x), y, rtol=1e-14, atol=1e-14)
    # a straight line is reproduced exactly, inside the knots and outside
    x = np.linspace(0, 3, 9)
    np.testing.assert_allclose(f(x, 2 * x - 1, [-1.0, 0.37, 5.0]), [-3.0, -0.26, 9.0], rtol=1e-14, atol=1e-14)
    with pytest.raises((ValueError, A.AusaxsError)):
        f([0, 1, 2], [0, 1, 2], [0.5])


def test_golden_section_closed_form():
    for c in (0.0, 3.3, 9.99, 7.123456):
        v, half, evals = O.golden(lambda x: (x - c) ** 2, 0, 10)
        assert abs(v - c) < 1e-4 and half < 0.5e-4
        # 2 starting points, n - 1 steps, the centre: n = ceil(log(1e-5) / log(1/phi)) = 24
        assert evals == 26
    v, half, evals = O.golden(lambda x: abs(x - 1.0), 0.99995, 1.00004)  # already below tol
    assert evals == 1 and half == pytest.approx(0.45e-4)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def simple_case():
    rng = np.random.default_rng(4)
    atoms = A.synth_globule(21, 900, (14, 14, 14), 2.0)
    atoms[:, 3] = rng.choice([6.0, 7.0, 8.0], len(atoms)).astype(F32)
    waters = A.synth_shell(22, 250, (14, 14, 14), 3.0)
    waters[:, 3] = 1.0
    return atoms, waters


def oracle_chi2(I_model_fn, q_model, q, y, err, params):
    out = []
    for p in params:
        m = O.spline(q_model, I_model_fn(p), q)
        out.append(O.lsq(m, y, err))
    return np.array(out)


@pytest.fixture(scope="module")
def simple_hist():
    atoms, waters = simple_case()
    s = A.default_settings(debye_terms=0)
    h = A.calculate_all(A.MODE_SIMPLE, atoms, waters, settings=s)
    ref = O.hist_simple(atoms, waters)
    return h, ref


def synthetic_data(h, q, cw, a=2.5e-3, b=0.4):
    h.set_params(cw=cw)
    I = a * h.debye_transform(q) + b
    h.set_params()
    return I, 0.01 * I


@pytest.mark.gpu
def test_fitter_batch_matches_oracle(simple_hist):
    h, ref = simple_hist
    q = np.linspace(0.01, 0.45, 90)
    y, err = synthetic_data(h, q, 1.7)
    f = A.Fitter(h, q, y, err)
    assert f.n_params == 1
    params = np.array([0.0, 0.5, 1.0, 2.5, 6.0])
    got = f.evaluate(params[:, None])
    qm = A.default_q_axis()
    exp = oracle_chi2(lambda p: O.debye_simple(ref, qm, 0.25, cw=p), qm, q, y, err, params)
    np.testing.assert_allclose(got[:, 0], exp[:, 0], rtol=1e-6)
    # b = weighted mean of y - a * model: its error scales with the data, not with b
    np.testing.assert_allclose(got[:, 1], exp[:, 1], rtol=1e-6, atol=1e-8 * np.max(np.abs(y)))
    np.testing.assert_allclose(got[:, 4], exp[:, 4], rtol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [A.MINI_GOLDEN, A.MINI_SCAN])
def test_fit_recovers_hydration(simple_hist, alg):
    h, ref = simple_hist
    q = np.linspace(0.01, 0.45, 90)
    y, err = synthetic_data(h, q, 1.7)
    r = A.Fitter(h, q, y, err).fit(alg)
    assert abs(r["params"][0] - 1.7) < 2e-3
    # the data were made at the data q values, the model is splined onto them:
    # b absorbs the interpolation error, small next to the data
    assert r["a"] == pytest.approx(2.5e-3, rel=1e-3) and abs(r["b"] - 0.4) < 1e-6 * np.max(y)
    assert r["dof"] == len(q) - 3 and r["chi2"] < 1e-3
    np.testing.assert_allclose(r["I_fit"], y, rtol=1e-4)
    if alg == A.MINI_GOLDEN:
        assert r["fevals"] == 26
    # the oracle's golden search on the oracle's chi2 lands on the same value
    qm = A.default_q_axis()
    v, _, _ = O.golden(lambda p: oracle_chi2(lambda c: O.debye_simple(ref, qm, 0.25, cw=c), qm, q, y, err,
                                             [p])[0, 4], 0, 10)
    assert abs(v - r["params"][0]) < 2e-3


@pytest.mark.gpu
def test_fit_ff_explicit_hydration():
    atoms = A.synth_globule(31, 1500, (16, 16, 16), 2.0)
    waters = A.synth_shell(32, 400, (16, 16, 16), 3.0)
    h = A.calculate_all(A.MODE_FFEXPLICIT, atoms, waters, settings=A.default_settings(debye_terms=0))
    q = np.linspace(0.02, 0.48, 70)
    y, err = synthetic_data(h, q, 0.6, a=1e-2, b=0.0)
    f = A.Fitter(h, q, y, err)
    r = f.fit(A.MINI_GOLDEN)
    assert abs(r["params"][0] - 0.6) < 2e-3 and r["chi2"] < 1e-3
    # histograms with excluded-volume terms accept the other parameters, one at a time
    f2 = A.Fitter(h, q, y, err, enabled=A.FIT_EXCLUDED_VOLUME)
    assert f2.n_params == 1
    assert f2.evaluate([[1.0], [1.05]]).shape == (2, 5)
    with pytest.raises(A.AusaxsError):
        A.Fitter(h, q, y, err, enabled=A.FIT_HYDRATION | A.FIT_EXCLUDED_VOLUME).fit()


@pytest.mark.gpu
def test_fitter_validation(simple_hist):
    h, _ = simple_hist
    q = np.linspace(0.01, 0.45, 30)
    y = np.ones(30)
    # a simple histogram keeps the hydration parameter only (validate_model)
    f = A.Fitter(h, q, y, enabled=A.FIT_HYDRATION | A.FIT_EXCLUDED_VOLUME | A.FIT_ATOMIC_DW)
    assert f.enabled == A.FIT_HYDRATION and f.n_params == 1
    with pytest.raises(A.AusaxsError):
        A.Fitter(h, q[:2], y[:2])
    with pytest.raises(A.AusaxsError):
        A.Fitter(h, q, y, q_model=q[:3])
    with pytest.raises(A.AusaxsError):
        A.Fitter(h, q, y).fit(algorithm=1)
