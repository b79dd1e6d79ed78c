"""Pins of the CPU oracle (no GPU).

Each test checks the oracle against something it does not restate: an
independent step-by-step IEEE emulation, a literal simulation of the
reference's loop structure, closed forms from the reference's own analytic
tests, physical invariants, or the reference's golden WAXS curve.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from conftest import compare_hist_approx, cube_atoms

F32 = np.float32


# --------------------------------------------------------------------------
# binning rules: independent IEEE emulation (numpy float32 ops + exact fma)
# --------------------------------------------------------------------------
def f32_round(fr: Fraction) -> np.float32:
    """Correctly rounded float32 of an exact rational (ties to even)."""
    f = F32(float(fr))
    cands = [np.nextafter(f, F32(-np.inf)), f, np.nextafter(f, F32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - fr), int(np.array(c).view(np.int32)) & 1))
    return F32(best)


def fma32(a, b, c):
    return f32_round(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def emul_simd(a, b, inv_w):
    dx, dy, dz = F32(a[0]) - F32(b[0]), F32(a[1]) - F32(b[1]), F32(a[2]) - F32(b[2])
    d2 = (dx * dx + dy * dy) + dz * dz
    d = np.sqrt(d2)
    x = d * F32(inv_w)
    return float(d), int(np.rint(x))


def emul_scalar(a, b, inv_w):
    dx, dy, dz = F32(a[0]) - F32(b[0]), F32(a[1]) - F32(b[1]), F32(a[2]) - F32(b[2])
    d2 = fma32(dz, dz, fma32(dx, dx, dy * dy))
    d = np.sqrt(F32(d2))
    x = d * F32(inv_w)
    return float(d), int(np.trunc(x + F32(0.49999997)))


@pytest.mark.parametrize("bw", [0.25, 0.1, 1.0])
def test_rules_match_ieee_emulation(bw):
    rng = np.random.default_rng(11)
    pts = (rng.standard_normal((400, 2, 4)) * 30).astype(np.float32)
    inv_w = O.inv_width(bw)
    n_fma_diff = 0
    for a, b in pts:
        d0, k0 = O.bin_pair(a, b, bw, 0)
        d1, k1 = O.bin_pair(a, b, bw, 1)
        assert (d0, k0) == emul_simd(a, b, inv_w)
        assert (d1, k1) == emul_scalar(a, b, inv_w)
        n_fma_diff += d0 != d1
    # the FMA-contracted squared distance differs from the separately rounded one often
    assert n_fma_diff > 10


def test_rules_ties_by_hand():
    # d = 0.625 (exact square, exact sqrt): x = 2.5 at 0.25 A bins.
    # SIMD (cvtps2dq, half to even) -> 2; scalar trunc(2.5 + 0.49999997f) = trunc(3.0f) -> 3.
    a = np.zeros(4, np.float32)
    b = np.array([0.625, 0, 0, 0], np.float32)
    assert O.bin_pair(a, b, 0.25, 0) == (0.625, 2)
    assert O.bin_pair(a, b, 0.25, 1) == (0.625, 3)
    # x = 3.5: both give 4
    b[0] = 0.875
    assert O.bin_pair(a, b, 0.25, 0)[1] == 4
    assert O.bin_pair(a, b, 0.25, 1)[1] == 4
    # just below a tie: both round down
    b[0] = np.nextafter(np.float32(0.625), np.float32(0))
    assert O.bin_pair(a, b, 0.25, 0)[1] == 2
    assert O.bin_pair(a, b, 0.25, 1)[1] == 2


def test_rules_integer_lattice_closed_form():
    # integer displacements: d = sqrt(k); |4 sqrt(k) - (m + 1/2)| >= 1/(32 sqrt(k)) keeps
    # float32 far from any tie for k <= 4096, so both rules equal round(4 sqrt(k)).
    a = np.zeros(4, np.float32)
    for dx in range(0, 40, 3):
        for dy in range(0, 40, 7):
            for dz in range(0, 40, 5):
                k = dx * dx + dy * dy + dz * dz
                if k > 4096:
                    continue
                b = np.array([dx, dy, dz, 0], np.float32)
                expect = int(np.floor(4 * np.sqrt(k) + 0.5))
                assert O.bin_pair(a, b, 0.25, 0)[1] == expect
                assert O.bin_pair(a, b, 0.25, 1)[1] == expect


# --------------------------------------------------------------------------
# loop geometry: literal simulation of the reference's 16/8/4/1 blocking
# --------------------------------------------------------------------------
def simulate_self(n):
    rule = {}
    for i in range(n):
        j = i + 1
        for width in (16, 8, 4):
            while j + width - 1 < n:
                for t in range(width):
                    rule[(i, j + t)] = 0
                j += width
        while j < n:
            rule[(i, j)] = 1
            j += 1
    return rule


def simulate_cross(m):
    rule = {}
    j = 0
    for width in (16, 8, 4):
        while j + width - 1 < m:
            for t in range(width):
                rule[j + t] = 0
            j += width
    while j < m:
        rule[j] = 1
        j += 1
    return rule


def test_loop_geometry():
    for n in range(0, 70):
        for (i, j), r in simulate_self(n).items():
            assert O.lib.orc_rule_self(n, i, j) == r, (n, i, j)
    for m in range(0, 70):
        for j, r in simulate_cross(m).items():
            assert O.lib.orc_rule_cross(m, j) == r, (m, j)


def test_pair_hist_matches_brute_force():
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 3, 5, 17, 38):
        a = (rng.random((n, 4)) * 20).astype(np.float32)
        h = O.pair_hist(a, None, 0.25, 400, factor=2)
        rules = simulate_self(n)
        cnt = np.zeros(400, np.int64)
        dsum = np.zeros(400)
        for (i, j), r in rules.items():
            d, k = (emul_simd if r == 0 else emul_scalar)(a[i], a[j], 4.0)
            cnt[k] += 2
            dsum[k] += 2 * d
        assert np.array_equal(h["count"], cnt)
        np.testing.assert_allclose(h["bin_center"], dsum, rtol=1e-12)


# --------------------------------------------------------------------------
# manager conventions: the cube of histogram_manager.cpp:30-146
# --------------------------------------------------------------------------
def cube_ff(water=False):
    atoms = np.zeros((8, 4), np.float32)
    atoms[:, :3] = cube_atoms()
    atoms.view(np.int32)[:, 3] = 3  # C
    waters = None
    if water:
        waters = atoms.copy()
        waters.view(np.int32)[:, 3] = 1
    return atoms, waters


@pytest.mark.parametrize("water", [False, True])
def test_cube_ffexplicit_counts(water):
    atoms, waters = cube_ff(water)
    h = O.hist_ff(O.FFEXPLICIT, atoms, waters)
    # 8 self, 12 edges (2 A -> bin 8), 12 face diagonals (sqrt 8 -> bin 11), 4 body (sqrt 12 -> bin 14), factor 2
    expect = {0: 8, 8: 24, 11: 24, 14: 8}
    assert h["max_bin"] == 15
    aa = h["aa"][3, 3]
    for b in range(15):
        assert aa["value"][b] == expect.get(b, 0)
    assert np.all(h["aa"]["value"].sum() == 64)
    d = O.weighted_axis(h["p_tot"], 0.25)
    for b, dist in ((8, 2.0), (11, 8 ** 0.5), (14, 12 ** 0.5)):
        assert abs(d[b] - dist) < 1e-6
    if water:
        assert list(h["aw"][3]["value"][[0, 8, 11, 14]]) == [8, 24, 24, 8]
        assert list(h["ww"]["value"][[0, 8, 11, 14]]) == [8, 24, 24, 8]


def test_cube_ffavg_exv_rows():
    atoms, waters = cube_ff(True)
    z = 1.7
    h = O.hist_ff(O.FFAVG, atoms, waters, z_exv=z)
    aa = h["aa"]["value"]
    # [C][exv]: pairs only (no self term), scaled by Z; [exv][exv]: pairs + N self, scaled by Z^2
    assert np.allclose(aa[3, 0, [0, 8, 11, 14]], np.array([0, 24, 24, 8]) * z)
    assert np.allclose(aa[0, 0, [0, 8, 11, 14]], np.array([8, 24, 24, 8]) * z * z)
    assert np.allclose(h["aw"]["value"][0, [0, 8, 11, 14]], np.array([8, 24, 24, 8]) * z)
    # p_tot excludes exv rows
    assert list(h["p_tot"]["value"][[0, 8, 11, 14]]) == [24, 72, 72, 24]


def test_simple_self_terms_and_counts():
    rng = np.random.default_rng(5)
    a = (rng.random((37, 4)) * 10).astype(np.float32)
    a[:, 3] = rng.choice([1.0, 2.5, 6.0], 37).astype(np.float32)
    w = (rng.random((13, 4)) * 10).astype(np.float32)
    w[:, 3] = 1.0
    h = O.hist_simple(a, w, 0.25, 400)
    n, m = len(a), len(w)
    assert h["aa"]["count"].sum() == n * (n - 1) + int(np.sum(a[:, 3].astype(np.float64) ** 2))
    assert h["aw"]["count"].sum() == 2 * n * m
    # value: sum over pairs of 2 w_i w_j + sum w^2 == (sum w)^2
    assert np.isclose(h["aa"]["value"].sum(), a[:, 3].astype(np.float64).sum() ** 2)
    assert np.isclose(h["aw"]["value"].sum(), 2 * a[:, 3].astype(np.float64).sum() * m)


# --------------------------------------------------------------------------
# sinc / Debye
# --------------------------------------------------------------------------
def test_sinc_table_closed_form():
    q = np.array([0.0, 1e-5, 1e-3, 0.1, 1.0, 7.0])
    d = np.array([0.0, 0.5, 2.0, 30.0, 123.4])
    S = O.sinc_table(q, d)
    x = np.outer(q, d)
    ref = np.where(x < 1e-3, 1 - x ** 2 / 6 + x ** 4 / 120, np.sinc(x / np.pi))
    np.testing.assert_allclose(S, ref, rtol=1e-13, atol=1e-15)
    assert np.all(S[:, 0] == 1.0)


def sinc(x):
    return np.sinc(np.asarray(x) / np.pi)


def test_cube_ffavg_intensity_closed_form():
    # composite_distance_histogram_ff_avg.cpp:23-86 "no water": 8 corners + centre, Z = 1, unweighted bins
    atoms = np.zeros((9, 4), np.float32)
    atoms[:8, :3] = cube_atoms()
    atoms.view(np.int32)[:, 3] = 3
    h = O.hist_ff(O.FFAVG, atoms, None, z_exv=1.0)
    q = 1e-4 + np.arange(200) * (1 - 1e-4) / 200
    I, _ = O.debye_ff(O.FFAVG, h, q, 0.25, weighted=False)
    w = 0.25
    d = [round(np.sqrt(3) / w) * w, round(2 / w) * w, round(np.sqrt(8) / w) * w, round(np.sqrt(12) / w) * w]
    pairs = 16 * sinc(q * d[0]) + 24 * sinc(q * d[1]) + 24 * sinc(q * d[2]) + 8 * sinc(q * d[3])
    aasum, axsum = 9 + pairs, pairs
    fc = np.array([O.ff_atomic(3, x) for x in q])
    fx = np.array([O.ff_atomic(0, x) for x in q])
    expect = aasum * fc ** 2 - 2 * axsum * fc * fx + aasum * fx ** 2
    np.testing.assert_allclose(I, expect, rtol=1e-10)


def test_cube_ffexplicit_intensity_closed_form():
    # explicit exv: every atom scatters with (f - f_exv); I = sum_ij (f - f_x)^2 sinc(q r_ij)
    atoms = np.zeros((9, 4), np.float32)
    atoms[:8, :3] = cube_atoms()
    atoms.view(np.int32)[:, 3] = 3
    h = O.hist_ff(O.FFEXPLICIT, atoms, None)
    q = np.linspace(0.0, 1.0, 77)
    I, _ = O.debye_ff(O.FFEXPLICIT, h, q, 0.25, weighted=True)
    d = [np.sqrt(3), 2, np.sqrt(8), np.sqrt(12)]
    aasum = 9 + 16 * sinc(q * d[0]) + 24 * sinc(q * d[1]) + 24 * sinc(q * d[2]) + 8 * sinc(q * d[3])
    f = np.array([O.ff_atomic(3, x) - O.ff_exv(3, x) for x in q])
    # float32 distances: ~3e-8 relative in the weighted bin centres
    np.testing.assert_allclose(I, aasum * f ** 2, rtol=1e-7)


def test_form_factors_electron_counts():
    # f(0) = number of electrons (5-Gaussian fits), excluded volume normalised to 1
    # (NH3 maps to the NH3+ group: 7 + 3 - 1 electrons)
    Z = {2: 1, 3: 6, 4: 7, 5: 8, 6: 9, 7: 7, 8: 8, 9: 9, 10: 9, 11: 8, 12: 16, 13: 17, 14: 18, 1: 9}
    for t, z in Z.items():
        assert abs(O.ff_atomic(t, 0.0) - z) < 0.02, t
    assert O.ff_atomic(0, 0.0) == 1.0
    for t in range(1, 15):
        qs = np.linspace(0, 3, 31)
        f = [O.ff_atomic(t, x) for x in qs]
        assert all(b <= a + 1e-12 for a, b in zip(f, f[1:])), t
        # exv form factor: V * 0.334 * exp(-V^(2/3)/(4 pi) q^2)
        v = O.ff_exv(t, 0.0) / 0.334
        assert np.isclose(O.ff_exv(t, 1.3), v * 0.334 * np.exp(-(v ** (2 / 3)) / (4 * np.pi) * 1.69))


def test_exact_debye_vs_binned():
    # composite_distance_histogram.cpp:277-352: binned Debye within 1 % of exact
    rng = np.random.default_rng(9)
    a = np.zeros((300, 4), np.float32)
    a[:, :3] = rng.standard_normal((300, 3)) * 8
    a[:, 3] = 1.0
    q = np.linspace(0.01, 0.5, 40)
    exact = O.exact_debye(a, q)
    h = O.hist_simple(a, None, 0.25, 800)
    d = O.weighted_axis(h["p_tot"], 0.25)
    binned = O.contract(h["p_tot"]["value"][None, :], O.sinc_table(q, d))[0] * np.exp(-q * q)
    np.testing.assert_allclose(binned, exact, rtol=1e-2)


def test_golden_carbon_sphere(golden_carbon):
    # feature/other/wide_angle_scattering.cpp:14-37: exv None, 0.1 A bins, 1 % (compare_hist_approx)
    xyz, q, I_ref = golden_carbon
    a = np.zeros((len(xyz), 4), np.float32)
    a[:, :3] = xyz.astype(np.float32)
    a[:, 3] = 1.0
    h = O.hist_simple(a, None, 0.1, 8000, nthreads=4)
    d = O.weighted_axis(h["p_tot"], 0.1)
    I = O.contract(h["p_tot"]["value"][None, :], O.sinc_table(q, d))[0]
    ok, i = compare_hist_approx(I, I_ref, 1e-6, 0.01)
    assert ok, (i, I[i], I_ref[i])


def test_lsq_closed_form():
    x = np.linspace(1, 10, 50)
    y = 2.5 * x - 3.0
    a, b, va, vb, chi2 = O.lsq(x, y, np.full(50, 0.3))
    assert np.isclose(a, 2.5) and np.isclose(b, -3.0) and chi2 < 1e-20
    # unit errors: var(a) = n / (n Sxx - Sx^2)
    _, _, va1, vb1, _ = O.lsq(x, y)
    n, sx, sxx = len(x), x.sum(), (x * x).sum()
    assert np.isclose(va1, n / (n * sxx - sx * sx)) and np.isclose(vb1, sxx / (n * sxx - sx * sx))


def test_lsq_matches_weighted_lstsq():
    # an independent solver: weighted linear least squares by numpy's lstsq;
    # chi2 is the minimised weighted residual norm
    rng = np.random.default_rng(2)
    x = rng.uniform(1, 100, 80)
    err = rng.uniform(0.5, 2.0, 80)
    y = 0.7 * x + 4.0 + rng.normal(0, 1, 80) * err
    A = np.stack([x, np.ones_like(x)], 1) / err[:, None]
    sol, res, *_ = np.linalg.lstsq(A, y / err, rcond=None)
    a, b, va, vb, chi2 = O.lsq(x, y, err)
    assert np.allclose([a, b], sol, rtol=1e-10)
    assert np.isclose(chi2, res[0], rtol=1e-9)
    cov = np.linalg.inv(A.T @ A)
    assert np.isclose(va, cov[0, 0], rtol=1e-9) and np.isclose(vb, cov[1, 1], rtol=1e-9)


def test_threads_do_not_change_counts():
    from conftest import globule

    atoms = globule(1500, 18.0, 1)
    h1 = O.hist_ff(O.FFEXPLICIT, atoms, None, nthreads=1)
    h4 = O.hist_ff(O.FFEXPLICIT, atoms, None, nthreads=4)
    assert np.array_equal(h1["aa"]["count"], h4["aa"]["count"])
    np.testing.assert_allclose(h1["aa"]["bin_center"], h4["aa"]["bin_center"], rtol=1e-12)


# --------------------------------------------------------------------------
# rigid bodies: PartialHistogramManagerMT
# --------------------------------------------------------------------------
def test_rigid_transform_closed_forms():
    rng = np.random.default_rng(3)
    xyz = rng.normal(size=(40, 3)) * 20
    w = rng.integers(1, 9, 40).astype(np.float32)
    ident = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0.0])
    p = O.rigid_transform(xyz, w, ident)
    assert np.array_equal(p[:, :3], xyz.astype(np.float32)) and np.array_equal(p[:, 3], w)
    # quarter turn about z plus a translation: (x, y, z) -> (-y + 1.5, x - 2, z + 0.25),
    # every product exact, one rounding per sum, then the float cast
    rz = np.array([0, -1, 0, 1, 0, 0, 0, 0, 1, 1.5, -2.0, 0.25])
    p = O.rigid_transform(xyz, w, rz)
    expect = np.stack([-xyz[:, 1] + 1.5, xyz[:, 0] - 2.0, xyz[:, 2] + 0.25], 1).astype(np.float32)
    assert np.array_equal(p[:, :3], expect)
    # a general rotation keeps distances up to the float casts
    th = 0.7
    R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    p = O.rigid_transform(xyz, w, np.concatenate([R.reshape(9), [3.0, 4.0, 5.0]]))
    d0 = np.linalg.norm(xyz[:, None] - xyz[None], axis=2)
    d1 = np.linalg.norm(p[:, None, :3].astype(np.float64) - p[None, :, :3], axis=2)
    assert np.max(np.abs(d0 - d1)) < 1e-4


def _partial_brute(self_bodies, bodies, waters, bw):
    """Pair-by-pair restatement with the pinned binning rules: self loops rule by
    (n-i-1) mod 4 geometry, cross loops by column index; add<2> per pair."""
    nb = 4000
    out = {k: np.zeros(nb, O.ENTRY) for k in ("aa", "aw", "ww")}

    def add(dst, d, b, w):
        out[dst]["count"][b] += 2
        out[dst]["value"][b] += 2.0 * float(np.float32(w))
        out[dst]["bin_center"][b] += 2.0 * d

    def self_loop(pts, dst):
        n = len(pts)
        for i in range(n):
            for j in range(i + 1, n):
                d, b = O.bin_pair(pts[i], pts[j], bw, O.lib.orc_rule_self(n, i, j))
                add(dst, d, b, np.float32(pts[i, 3]) * np.float32(pts[j, 3]))
        t = float(np.sum((pts[:, 3] * pts[:, 3]).astype(np.float64)))
        out[dst]["value"][0] += t
        out[dst]["count"][0] += int(t)

    def cross(rows, cols, dst):
        m = len(cols)
        for i in range(len(rows)):
            for j in range(m):
                d, b = O.bin_pair(rows[i], cols[j], bw, O.lib.orc_rule_cross(m, j))
                add(dst, d, b, np.float32(rows[i, 3]) * np.float32(cols[j, 3]))

    for i in range(len(bodies)):
        self_loop(self_bodies[i], "aa")
        for j in range(i):
            cross(bodies[j], bodies[i], "aa")  # calc_aa(i, j): rows over body j
        if len(waters):
            cross(waters, bodies[i], "aw")  # calc_aw(i): rows over the waters
    if len(waters):
        self_loop(waters, "ww")
    return out


def test_partial_manager_brute_force():
    rng = np.random.default_rng(17)
    sizes = [5, 6, 7, 9]
    T = []
    self_b, moved = [], []
    for k, n in enumerate(sizes):
        xyz = rng.normal(size=(n, 3)) * 4 + k * 6.0
        w = rng.integers(1, 5, n).astype(np.float32)
        self_b.append(O.rigid_transform(xyz, w, [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]))
        c, s = np.cos(0.3 * k), np.sin(0.3 * k)
        Tk = [c, -s, 0, s, c, 0, 0, 0, 1, 0.5 * k, -0.25, 1.0]
        moved.append(O.rigid_transform(xyz, w, Tk))
    waters = np.zeros((6, 4), np.float32)
    waters[:, :3] = rng.normal(size=(6, 3)) * 8
    waters[:, 3] = 1.0
    h = O.hist_partial(np.concatenate(self_b), np.concatenate(moved), sizes, waters, 0.25, 4000)
    ref = _partial_brute(self_b, moved, waters, 0.25)
    mb = h["max_bin"]
    for key in ("aa", "aw", "ww"):
        assert np.array_equal(h[key]["count"], ref[key]["count"][:mb]), key
        np.testing.assert_allclose(h[key]["value"], ref[key]["value"][:mb], rtol=1e-12)
        np.testing.assert_allclose(h[key]["bin_center"], ref[key]["bin_center"][:mb], rtol=1e-6, atol=1e-9)
    assert not np.any(ref["aa"]["count"][mb:])


def test_partial_equals_full_when_unmoved():
    # partial_histogram_manager.cpp:70-154: the partial manager reproduces the
    # full manager on the same molecule
    rng = np.random.default_rng(23)
    sizes = [37, 52, 61]
    pts = np.zeros((sum(sizes), 4), np.float32)
    pts[:, :3] = rng.normal(size=(sum(sizes), 3)) * 12
    pts[:, 3] = rng.integers(1, 9, sum(sizes))
    waters = np.zeros((23, 4), np.float32)
    waters[:, :3] = rng.normal(size=(23, 3)) * 16
    waters[:, 3] = 1.0
    hp = O.hist_partial(pts, pts, sizes, waters, 0.25, 8000)
    hf = O.hist_simple(pts, waters, 0.25, 8000)
    assert hp["max_bin"] == hf["max_bin"]
    for key in ("aa", "aw", "ww", "p_tot"):
        assert np.array_equal(hp[key]["count"], hf[key]["count"]), key
        np.testing.assert_allclose(hp[key]["value"], hf[key]["value"], rtol=1e-13)
        # the partial manager's inter-body blocks are cross loops, whose tail
        # columns use the scalar (FMA) rule where the full self loop used the
        # SIMD rule: distances agree to float rounding
        np.testing.assert_allclose(hp[key]["bin_center"], hf[key]["bin_center"], rtol=1e-6, atol=1e-9)


def test_debye_simple_golden(golden_carbon):
    xyz, q, I_ref = golden_carbon
    a = np.zeros((len(xyz), 4), np.float32)
    a[:, :3] = xyz.astype(np.float32)
    a[:, 3] = 1.0
    h = O.hist_simple(a, None, 0.1, 8000, nthreads=4)
    I = O.debye_simple(h, q, 0.1) * np.exp(q * q)
    ok, i = compare_hist_approx(I, I_ref, 1e-6, 0.01)
    assert ok, (i, I[i], I_ref[i])
