"""Grid excluded volume with a surface layer (HistogramManagerMTFFGridSurface,
HistogramManagerMTFFGridSurface.cpp:40-340; CompositeDistanceHistogramFFGridSurface.cpp:60-266;
RawGridWithSurfaceExv.cpp:14-127).

The dummies of the grid split into interior and surface ones by a radial
collision score; the surface dummies' form factor is scaled by
exv_factor(q, cx) = cx^3 exp(-rm^2/4 (cx^2 - 1) q^2) when the fit varies cx.

Pins of the oracle:
* the classification reproduces the reference's own known answers
  (tests/feature/grid/grid_surface_detection.cpp:40-187, unit surface thickness,
  atomic radius 0 and the minimum excluded-volume radius each case sets);
* interior and surface partition the plain grid's dummies, in order;
* the interior / surface / cross histograms add up bin for bin to the FFGrid
  oracle's histograms of all dummies, and at cx = 1 the two intensities agree;
* with cx != 1 the xx and wx profiles equal the exact Debye sums over the
  dummies weighted 1 (interior) and exv_factor (surface), to the binning error.
The product is then held to the oracle: the classification identical, counts
bit for bit, I(q) within 1e-4.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

C_TYPE = 3  # carbon

# (atoms, min_exv_radius, (interior, surface)): grid_surface_detection.cpp:40-187
CUBE2 = [(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)]
CUBE3 = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)]
KNOWN = [
    ("single with radius", [(0, 0, 0)], 3 ** 0.5 + 1e-3, (1, 26)),
    ("two atoms with radius", [(0, 0, 0), (1, 0, 0)], 2 ** 0.5 + 1e-3, (2, 26)),
    ("2x2x2 interior simple", CUBE2, 2.0, (32, 56)),
    ("2x2x2 interior advanced", CUBE2, 2 ** 0.5 + 1e-3, (8, 48)),
    ("3x3x3", CUBE3, 2.0, (81, 98)),
    ("larger radius, single", [(0, 0, 0)], 2.0, (1, 32)),
    ("much larger radius, single", [(0, 0, 0)], 3.0, (27, 96)),
    ("larger radius, larger cube", [(-2, 0, 0), (2, 0, 0), (0, -2, 0), (0, 2, 0)], 2.0, (21, 92)),
]


@pytest.mark.parametrize("name,atoms,rvol,expected", KNOWN, ids=[k[0] for k in KNOWN])
def test_oracle_classification_known_answers(name, atoms, rvol, expected):
    a = np.array(atoms, np.float64)
    xi, xs, _ = O.grid_exv_surface(a, np.full(len(a), C_TYPE), rvol=rvol, atomic_radius=0.0)
    assert (len(xi), len(xs)) == expected


def structure(seed=3, n=400, m=120, r=10.0):
    g = A.synth_globule(seed, n, (r, r, r), 2.0)
    w = A.synth_shell(seed + 1, m, (r, r, r), 3.0)[:, :3].astype(np.float64) if m else np.zeros((0, 3))
    return g[:, :3].astype(np.float64), g[:, 3].view(np.int32).copy(), w


@pytest.mark.parametrize("seed,n,m", [(1, 80, 0), (3, 400, 120), (5, 900, 300)])
def test_partition_and_product_classification(seed, n, m):
    xyz, ff, w = structure(seed, n, m, 7.0 + n / 100)
    xi, xs, vol = O.grid_exv_surface(xyz, ff, w)
    every, vol0 = O.grid_exv(xyz, ff, w)
    assert vol == vol0 and len(xi) + len(xs) == len(every) and len(xi) > 0 and len(xs) > 0
    # interior and surface keep the grid's lexicographic order
    keys = {tuple(p): k for k, p in enumerate(every)}
    for part in (xi, xs):
        idx = [keys[tuple(p)] for p in part]
        assert idx == sorted(idx)
    assert len({tuple(p) for p in xi} | {tuple(p) for p in xs}) == len(every)
    # the product's host-side classification is the oracle's, point for point
    pi, ps = A.grid_exv_surface(xyz, ff, w)
    assert np.array_equal(pi, xi) and np.array_equal(ps, xs)


def test_surface_histograms_add_up_to_ffgrid():
    xyz, ff, w = structure()
    s = O.hist_ffgrid_surface(xyz, ff, w)
    g = O.hist_ffgrid(xyz, ff, w)
    ni, ns, n, m = len(s["interior"]), len(s["surface"]), len(xyz), len(w)
    assert s["max_bin"] == g["max_bin"]
    c = lambda k: s[k]["count"].astype(np.int64)
    assert c("xx_interior").sum() == ni * ni and c("xx_surface").sum() == ns * ns
    assert c("xx_cross").sum() == 2 * ni * ns
    assert c("ax_interior").sum() == n * ni and c("ax_surface").sum() == n * ns
    assert c("wx_interior").sum() == m * ni and c("wx_surface").sum() == m * ns
    # union of interior and surface = all dummies: the split histograms sum to
    # FFGrid's, bin for bin for the dummy pairs (integer displacements: which
    # loop tail takes the scalar rule cannot move a bin); for atom and water
    # pairs a near-tie pair may move by one bin, so totals exactly, bins to a few
    xx = c("xx_interior") + c("xx_surface") + c("xx_cross")
    assert np.array_equal(xx, g["aa"]["count"][0, 0])
    ax = (c("ax_interior") + c("ax_surface"))[1:]
    ax_g = g["aa"]["count"][1:, 0].astype(np.int64)  # [0][0] is xx
    wx, wx_g = c("wx_interior") + c("wx_surface"), g["aw"]["count"][0].astype(np.int64)
    assert ax.sum() == ax_g.sum() and np.abs(ax - ax_g).sum() <= 4
    assert wx.sum() == wx_g.sum() and np.abs(wx - wx_g).sum() <= 4
    # the atom and water histograms are the same computation in both managers
    assert np.array_equal(s["aa"]["count"][1:, 1:], g["aa"]["count"][1:, 1:])
    assert np.array_equal(s["aw"]["count"][1:], g["aw"]["count"][1:])
    assert np.array_equal(s["ww"]["count"], g["ww"]["count"])
    assert np.array_equal(s["p_xx"]["count"], g["p_xx"]["count"])
    assert s["p_tot_ax"]["count"].sum() == g["p_tot_ax"]["count"].sum()
    # the bin-centre sums add the same terms in another order
    np.testing.assert_allclose(s["d_ax"], g["d_ax"], rtol=1e-9)
    np.testing.assert_allclose(s["d_xx"], g["d_xx"], rtol=1e-9)
    # at cx = 1 the surface scaling is 1: the FFGrid intensity
    q = np.linspace(0.0, 0.5, 80)
    I_s, _ = O.debye_ffgrid_surface(s, q, 0.25, cw=1.2, crho=0.95)
    I_g, _ = O.debye_ffgrid(g, q, 0.25, cw=1.2, crho=0.95)
    np.testing.assert_allclose(I_s, I_g, rtol=1e-9)


@pytest.mark.parametrize("cx", [0.6, 1.45])
def test_surface_scaling_against_exact_debye(cx):
    # xx and wx profiles at cx against sum_ij w_i w_j sinc(q r_ij) with w = 1
    # (interior), exv_factor(q, cx) (surface); binning at 0.25 A on weighted
    # centres leaves ~1e-4 relative error at these q
    xyz, ff, w = structure(7, 120, 60, 6.0)
    s = O.hist_ffgrid_surface(xyz, ff, w)
    q = np.linspace(0.01, 0.3, 12)
    _, prof = O.debye_ffgrid_surface(s, q, 0.25, cx=cx)
    fx = 0.334 * np.exp(-q * q / (4 * np.pi))
    fw = np.array([O.ff_atomic(1, v) for v in q])
    c = O.exv_factor(q, cx)
    wat = np.zeros((len(w), 4), np.float32)
    wat[:, :3] = w
    wat[:, 3] = 1
    xx_exact, wx_exact = np.zeros(len(q)), np.zeros(len(q))
    for k in range(len(q)):
        d = np.zeros((len(s["interior"]) + len(s["surface"]), 4), np.float32)
        d[:, :3] = np.vstack([s["interior"], s["surface"]])
        d[:, 3] = 1.0
        d[len(s["interior"]):, 3] = c[k]
        # exact_debye carries a unit form factor exp(-q^2 / 2) per point
        unit = np.exp(q[k] * q[k])
        e_d = O.exact_debye(d, q[k:k + 1])[0] * unit
        e_w = O.exact_debye(wat, q[k:k + 1])[0] * unit
        e_dw = O.exact_debye(np.vstack([d, wat]), q[k:k + 1])[0] * unit
        xx_exact[k] = e_d
        wx_exact[k] = e_dw - e_d - e_w  # 2 sum_{d,w} w_d sinc
    # to the binning error, a few 1e-4 of the profile's peak (exchanging c and
    # c^2, or dropping the cross term, moves them by tens of percent)
    for got, ref in ((prof[3] / (fx * fx), xx_exact), (prof[4] / (fw * fx), wx_exact)):
        assert np.max(np.abs(got - ref)) <= 3e-4 * np.max(np.abs(ref))


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def relerr(x, ref):
    return float(np.max(np.abs(x - ref) / np.abs(ref)))


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,m,terms", [(3, 400, 120, 3), (9, 1500, 400, 3), (9, 1500, 400, 0)])
def test_ffgrid_surface_parity(seed, n, m, terms):
    xyz, ff, w = structure(seed, n, m, 7.0 + n / 100)
    h = A.GridSurfaceStructure(xyz, ff, w, settings=A.default_settings(debye_terms=terms)).calculate()
    ref = O.hist_ffgrid_surface(xyz, ff, w, nthreads=4)
    assert h.mode == A.MODE_FFGRID_SURFACE and h.max_bin == ref["max_bin"]
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT_AX, "p_tot_ax"), (A.XX, "p_xx"),
                       (A.XX_INTERIOR, "xx_interior"), (A.XX_SURFACE, "xx_surface"), (A.XX_CROSS, "xx_cross"),
                       (A.WX_INTERIOR, "wx_interior"), (A.WX_SURFACE, "wx_surface"),
                       (A.AX_INTERIOR, "ax_interior"), (A.AX_SURFACE, "ax_surface")):
        v, c, _ = h.get(which)
        assert np.array_equal(c, ref[key]["count"]), key
        np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-14, err_msg=key)
    q = np.linspace(0.0, 0.5, 120)
    for params in (dict(), dict(cx=0.7), dict(cw=1.3, cx=1.4, crho=1.05)):
        I_ref, P_ref = O.debye_ffgrid_surface(ref, q, 0.25, **params)
        h.set_params(**params)
        I, P = h.debye_transform(q, profiles=True)
        # The 1500-atom globule's exv nearly cancels its atoms near q = 0.03:
        # I there is 3e-6 of the largest profile. The three-term bf16 split
        # contraction carries about 2^-24 of each profile's magnitude, so its
        # bound there is absolute in the profiles' sum; the FP64 contraction
        # (terms = 0) holds 1e-4 of I at every q.
        scale = np.sum(np.abs(P_ref), axis=0)
        err = relerr(I, I_ref)
        err_p = float(np.max(np.abs(I - I_ref) / scale))
        print(f"ffgrid-surface terms={terms} {params} I(q) max rel err {err:.3e}, of the profile sum {err_p:.3e}")
        assert np.all(np.abs(I - I_ref) <= 1e-4 * np.abs(I_ref) + (2e-8 if terms else 0.0) * scale)
        for k in range(6):
            scale = np.max(np.abs(P_ref[k]))
            assert np.max(np.abs(P[k] - P_ref[k])) <= 1e-4 * scale, k


@pytest.mark.gpu
def test_fit_surface_scaling_outside_the_plain_limits():
    # FFGridSurface widens the excluded-volume limits to [0, 2]
    # (CompositeDistanceHistogramFFGridSurface.cpp:47-49): data made at cx = 1.3
    xyz, ff, w = structure(11, 700, 200, 11.0)
    h = A.GridSurfaceStructure(xyz, ff, w).calculate()
    q = np.linspace(0.02, 0.45, 80)
    h.set_params(cx=1.3)
    y = 3e-3 * h.debye_transform(q) + 0.1
    h.set_params()
    r = A.Fitter(h, q, y, 0.01 * np.abs(y), enabled=A.FIT_EXCLUDED_VOLUME).fit(A.MINI_GOLDEN)
    assert abs(r["params"][0] - 1.3) < 2e-3 and r["chi2"] < 1e-2


@pytest.mark.gpu
def test_surface_large_invariants():
    # 20k-atom globule (radius 45 A) + 6k waters, dummy pairs by lattice
    # correlation (automatic at this size): exact pair-count identities, and
    # the split histograms add up to the plain grid's bin for bin
    g = A.synth_globule(2, 20000, (45, 45, 45), 2.0)
    xyz, ff = g[:, :3].astype(np.float64), g.view(np.int32)[:, 3].copy()
    w = A.synth_shell(3, 6000, (45, 45, 45), 3.0)[:, :3].astype(np.float64)
    xi, xs = A.grid_exv_surface(xyz, ff, w)
    ni, ns, n, m = len(xi), len(xs), len(xyz), len(w)
    st = A.GridSurfaceStructure(xyz, ff, w)
    st.set_timing(True)
    h = st.calculate()
    assert st.last_timing()["lattice_pairs"] > 0
    c = lambda which: h.get(which)[1].astype(np.int64)
    assert c(A.XX_INTERIOR).sum() == ni * ni and c(A.XX_SURFACE).sum() == ns * ns
    assert c(A.XX_CROSS).sum() == 2 * ni * ns
    assert c(A.AX_INTERIOR).sum() == n * ni and c(A.AX_SURFACE).sum() == n * ns
    assert c(A.WX_INTERIOR).sum() == m * ni and c(A.WX_SURFACE).sum() == m * ns
    plain = A.GridStructure(xyz, ff, w).calculate()
    assert plain.max_bin == h.max_bin
    # dummy pairs: integer displacements, so the split into loops (which
    # pairs take the reference's scalar tail rule) cannot move a pair's bin
    xx = c(A.XX_INTERIOR) + c(A.XX_SURFACE) + c(A.XX_CROSS)
    assert np.array_equal(xx, plain.get(A.XX)[1].astype(np.int64))
    # water-dummy pairs: the tail rule can move a near-tie pair by one bin, so
    # the totals agree exactly and the bins to a handful of pairs
    wx = c(A.WX_INTERIOR) + c(A.WX_SURFACE)
    wx_plain = plain.get(A.AW)[1][0].astype(np.int64)
    assert wx.sum() == wx_plain.sum() and np.abs(wx - wx_plain).sum() <= 8
