"""Grid-surface excluded volume with a surface thickness other than one cell.

settings::grid::exv::surface_thickness (GridSettings.cpp:17, default 1 A)
selects RawGridWithSurfaceExv's non-unity branch (RawGridWithSurfaceExv.cpp:
137-190, 196-204) when round(thickness / cell_width) != 1: the first pass marks
the surface dummies, every dummy within expand = int(round(thickness) * 0.5)
cells of one (squared integer distance <= expand^2) joins the surface, and a
last pass collects both sets in voxel order.

The oracle's restatement is pinned here by the definition, evaluated by brute
force over the oracle's own one-cell split: the thick surface is exactly the
set of dummies within expand cells of a one-cell surface dummy, and both parts
keep the grid's lexicographic order. The product's host-side split is then
compared with the oracle point for point, and (GPU) its histograms bin for bin.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O


def structure(seed, n, m, r):
    g = A.synth_globule(seed, n, (r, r, r), 2.0)
    xyz, ff = g[:, :3].astype(np.float64), g.view(np.int32)[:, 3].copy()
    w = A.synth_shell(seed + 1, m, (r, r, r), 3.0)[:, :3].astype(np.float64) if m else np.zeros((0, 3))
    return xyz, ff, w


def expected_split(every, surface1, expand):
    """Brute force: a dummy is on the thick surface when a one-cell surface
    dummy lies within squared integer distance expand^2 of it."""
    s1 = np.asarray(surface1)
    d2 = ((every[:, None, :] - s1[None, :, :]) ** 2).sum(-1)
    on = (d2 <= expand * expand + 1e-9).any(axis=1)
    return every[~on], every[on]


@pytest.mark.parametrize("seed,n,m", [(3, 300, 80), (5, 700, 200)])
@pytest.mark.parametrize("thickness", [0.4, 1.0, 2.0, 3.0, 4.0, 5.0])
def test_oracle_thick_surface_by_definition(seed, n, m, thickness):
    xyz, ff, w = structure(seed, n, m, 6.0 + n / 100)
    every, _ = O.grid_exv(xyz, ff, w)
    xi1, xs1, _ = O.grid_exv_surface(xyz, ff, w)
    xi, xs, _ = O.grid_exv_surface(xyz, ff, w, surface_thickness=thickness)
    expand = int(round(thickness) * 0.5)
    ei, es = expected_split(every, xs1, expand)
    assert np.array_equal(xi, ei) and np.array_equal(xs, es)
    if round(thickness) == 1 or expand == 0:
        # the unity branch, or the non-unity branch with a ball of one cell
        assert np.array_equal(xs, xs1) and np.array_equal(xi, xi1)
    else:
        assert len(xs) > len(xs1) and len(xi) + len(xs) == len(every)


@pytest.mark.parametrize("thickness", [0.4, 2.0, 3.0, 4.0, 6.0])
def test_product_thick_surface_matches_oracle(thickness):
    xyz, ff, w = structure(7, 900, 250, 9.0)
    xi, xs, _ = O.grid_exv_surface(xyz, ff, w, surface_thickness=thickness)
    pi, ps = A.grid_exv_surface(xyz, ff, w, surface_thickness=thickness)
    assert np.array_equal(pi, xi) and np.array_equal(ps, xs)


def test_thickness_out_of_range_is_an_error():
    xyz, ff, w = structure(3, 100, 0, 5.0)
    for bad in (-1.0, float("nan"), 100.0):
        with pytest.raises(A.AusaxsError, match="surface_thickness"):
            A.grid_exv_surface(xyz, ff, w, surface_thickness=bad)


@pytest.mark.gpu
@pytest.mark.parametrize("thickness", [2.0, 4.0])
def test_ffgrid_surface_thick_parity(thickness):
    xyz, ff, w = structure(9, 1200, 300, 10.0)
    st = A.GridSurfaceStructure(xyz, ff, w, settings=A.default_settings(surface_thickness=thickness))
    h = st.calculate()
    ref = O.hist_ffgrid_surface(xyz, ff, w, nthreads=4, surface_thickness=thickness)
    assert h.max_bin == ref["max_bin"]
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.XX, "p_xx"),
                       (A.XX_INTERIOR, "xx_interior"), (A.XX_SURFACE, "xx_surface"), (A.XX_CROSS, "xx_cross"),
                       (A.WX_INTERIOR, "wx_interior"), (A.WX_SURFACE, "wx_surface"),
                       (A.AX_INTERIOR, "ax_interior"), (A.AX_SURFACE, "ax_surface")):
        v, c, _ = h.get(which)
        assert np.array_equal(c, ref[key]["count"]), key
        np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-14, err_msg=key)
    q = np.linspace(0.0, 0.5, 100)
    for params in (dict(), dict(cx=0.8)):
        I_ref, P_ref = O.debye_ffgrid_surface(ref, q, 0.25, **params)
        h.set_params(**params)
        I = h.debye_transform(q)
        scale = np.sum(np.abs(P_ref), axis=0)
        assert np.all(np.abs(I - I_ref) <= 1e-4 * np.abs(I_ref) + 2e-8 * scale)
    # a thicker surface moves dummies from the interior to the surface set
    h1 = A.GridSurfaceStructure(xyz, ff, w).calculate()
    n_s = lambda hh: int(hh.get(A.XX_SURFACE)[1].sum())
    assert n_s(h) > n_s(h1)
