"""Radial hydration shell (RadialHydration, RadialHydration.cpp:40-206; the grid
of GridBasedHydration::hydrate, GridBasedHydration.cpp:36-79; NoCulling).

The placement is host code in both the product and the oracle. The
reference's known-answer case (tests/feature/grid/grid.cpp:707-748: one carbon
atom at the origin of a grid spanning -10..10 A, atomic and hydration radius 3
A, shell correction 0, zero noise) requires waters at the six axis sites 6 A
out; both reproduce it, and variants of it whose outcome follows from the
geometry alone. Its zero-noise consistency test
(tests/feature/hydrate/radial_hydration.cpp) requires identical waters on
every call. Beyond those the oracle is pinned by what the procedure's
definition implies, checked on its output with plain geometry:
every water sits at vdw + 1.36 A from an atom along one of the 26 directions;
no water's voxel is an atom centre, lies inside an atom's van der Waals
sphere or repeats another water's voxel, and one within 1.71 A of an earlier
water lies in an atom's 2.15 A volume shell; and bodies processed one after
the other on the shared grid give the one-body result split in order. The
product is then held to the oracle point for point, with and without noise.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

VDW = {2: 1.1, 3: 1.96, 4: 1.96, 5: 1.96, 6: 1.96, 7: 1.79, 8: 1.79, 9: 1.79, 10: 1.79, 1: 1.71, 11: 1.71,
       12: 2.06, 13: 2.06, 14: 1.88}


def structure(seed, n):
    r = 20.0 * (n / 2000.0) ** (1 / 3)
    g = A.synth_globule(seed, n, (r, r, r), 2.0)
    return g[:, :3].astype(np.float64), g[:, 3].view(np.int32).copy()


def directions():
    """The 26 directions: faces, edges (45 degrees) and the 8 octant diagonals
    at theta = phi = 45 degrees, from the angles themselves."""
    s, c = np.sin(np.pi / 4), np.cos(np.pi / 4)
    base = {(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (s, 0.0, c), (0.0, s, c), (s, s, 0.0),
            (c * s, s * s, c)}
    out = set()
    for b in base:
        for sx in (1, -1):
            for sy in (1, -1):
                for sz in (1, -1):
                    out.add((round(sx * b[0], 12) + 0.0, round(sy * b[1], 12) + 0.0, round(sz * b[2], 12) + 0.0))
    return np.array(sorted(out))


def test_directions_are_the_26_of_the_octant_construction():
    assert len(directions()) == 26


SIX_SITES = {(0, 0, 6), (0, 0, -6), (6, 0, 0), (-6, 0, 0), (0, 6, 0), (0, -6, 0)}
LIMITS = ((-10, -10, -10), (10, 10, 10))


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_reference_known_answer_six_axis_sites(impl):
    # tests/feature/grid/grid.cpp:707-748 (RadialStrategy): GridDebug with
    # atomic and hydration radius 3 on Limit3D(-10, 10, ...) at 1 A cells,
    # shell_correction 0, zero noise; REQUIRE(6 <= waters) and exactly the six
    # axis sites among them
    f = O.hydrate_radial if impl == "oracle" else A.hydrate
    w, per = f(np.zeros((1, 3)), np.array([3], np.int32), atomic_radius=3, hydration_radius=3, shell_correction=0,
               grid_limits=LIMITS)
    assert len(w) >= 6 and per.tolist() == [len(w)]
    sites = [tuple(int(v) for v in np.rint(x)) for x in w if np.allclose(x, np.rint(x), atol=0, rtol=0)]
    assert sum(1 for t in sites if t in SIX_SITES) == 6
    # here no other candidate passes: the 20 off-axis directions put the
    # water's 1-shell-radius probes back onto the atom's 3 A sphere
    assert len(w) == 6


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_known_answer_variants_follow_the_geometry(impl):
    f = O.hydrate_radial if impl == "oracle" else A.hydrate
    a, t = np.zeros((1, 3)), np.array([3], np.int32)
    dirs = directions()
    for corr in (1.0, -0.5):
        # every water sits at ra + rh + shell_correction along one of the 26
        # directions, inside the grid
        w, _ = f(a, t, atomic_radius=3, hydration_radius=3, shell_correction=corr, grid_limits=LIMITS)
        assert len(w) > 0
        reff = 6.0 + corr
        assert np.allclose(np.linalg.norm(w, axis=1), reff, atol=1e-12)
        assert all(np.min(np.linalg.norm(dirs - x / reff, axis=1)) < 1e-12 for x in w)
    # the atom and the grid moved together by whole cells move every water with them
    w0, _ = f(a, t, atomic_radius=3, hydration_radius=3, shell_correction=0, grid_limits=LIMITS)
    shift = np.array([1.0, -2.0, 3.0])
    moved = (tuple(np.add(LIMITS[0], shift)), tuple(np.add(LIMITS[1], shift)))
    w1, _ = f(a + shift, t, atomic_radius=3, hydration_radius=3, shell_correction=0, grid_limits=moved)
    assert np.array_equal(w1, w0 + shift)
    # a candidate whose voxel falls outside the grid is never placed
    small = ((-6, -6, -6), (6, 6, 6))  # bins 0..11 hold -6..5: +6 rounds to bin 12
    w2, _ = f(a, t, atomic_radius=3, hydration_radius=3, shell_correction=0, grid_limits=small)
    assert np.all(np.rint(w2) < 6)
    # an atom outside the explicit grid is an error
    with pytest.raises((ValueError, A.AusaxsError)):
        f(a + 20.0, t, atomic_radius=3, hydration_radius=3, shell_correction=0, grid_limits=LIMITS)


@pytest.mark.parametrize("seed,n", [(1, 300), (2, 2000)])
def test_oracle_definition_invariants(seed, n):
    xyz, ff = structure(seed, n)
    w, per = O.hydrate_radial(xyz, ff)
    assert len(w) > 0 and per.sum() == len(w)
    # the reference's consistency test: zero noise gives identical waters every call
    w2, _ = O.hydrate_radial(xyz, ff)
    assert np.array_equal(w, w2)
    reff = np.array([VDW[int(f)] for f in ff]) + 1.71 - 0.35
    dirs = directions()
    for p in w:
        d = p - xyz
        dist = np.sqrt((d * d).sum(1))
        k = np.flatnonzero(np.abs(dist - reff) < 1e-9)
        assert len(k), p
        u = d[k] / reff[k][:, None]
        assert np.min(np.abs(u[:, None, :] - dirs[None, :, :]).max(-1)) < 1e-9
    vox = np.round(w)  # the grid origin is an integer and the cell 1 A
    assert len({tuple(v) for v in vox}) == len(vox)
    atom_vox = {tuple(v) for v in np.round(xyz)}
    assert not atom_vox & {tuple(v) for v in vox}
    rv = np.array([VDW[int(f)] for f in ff])
    for v in vox:
        d2 = ((xyz - v) ** 2).sum(1)
        assert np.all(d2 > rv ** 2)
    # an earlier water claims the empty voxels within 1.71 A of it; a voxel of
    # an atom's 2.15 A volume shell stays open
    for i in range(1, len(w)):
        d2 = ((w[:i] - vox[i]) ** 2).sum(1)
        if np.any(d2 <= 1.71 ** 2):
            assert np.min(((xyz - vox[i]) ** 2).sum(1)) <= 2.15 ** 2


def test_bodies_in_order_split_the_one_body_result():
    xyz, ff = structure(3, 1500)
    w1, per1 = O.hydrate_radial(xyz, ff)
    w3, per3 = O.hydrate_radial(xyz, ff, bodies=[400, 700, 400])
    assert np.array_equal(w1, w3) and per3.sum() == per1[0] and len(per3) == 3


@pytest.mark.parametrize("seed,n,bodies", [(1, 300, None), (2, 2000, [500, 1500]), (4, 6000, [1000, 2000, 3000])])
def test_product_matches_oracle(seed, n, bodies):
    xyz, ff = structure(seed, n)
    w1, p1 = A.hydrate(xyz, ff, bodies=bodies)
    w2, p2 = O.hydrate_radial(xyz, ff, bodies=bodies)
    assert np.array_equal(w1, w2) and np.array_equal(p1, p2)
    noise = np.random.default_rng(seed).normal(0.0, 0.75, (n, 26, 3))
    w3, _ = A.hydrate(xyz, ff, bodies=bodies, noise=noise)
    w4, _ = O.hydrate_radial(xyz, ff, bodies=bodies, noise=noise)
    assert np.array_equal(w3, w4)
    # the seeded form draws the same distribution
    w5, _ = A.hydrate(xyz, ff, bodies=bodies, noise=seed)
    assert np.array_equal(w5, w3)


def test_input_errors():
    xyz, ff = structure(5, 50)
    with pytest.raises(A.AusaxsError):
        A.hydrate(xyz, ff, bodies=[10, 10])
    bad = ff.copy()
    bad[3] = 0
    with pytest.raises(A.AusaxsError):
        A.hydrate(xyz, bad)
    w, per = A.hydrate(np.zeros((0, 3)), np.zeros(0, np.int32))
    assert len(w) == 0


@pytest.mark.gpu
def test_hydrated_structure_through_the_grid_path():
    # atoms -> hydration shell -> FFGrid histograms and I(q), against the oracle
    xyz, ff = structure(6, 1500)
    w, _ = A.hydrate(xyz, ff, noise=6)
    h = A.GridStructure(xyz, ff, w).calculate()
    ref = O.hist_ffgrid(xyz, ff, w, nthreads=4)
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww")):
        assert np.array_equal(h.get(which)[1], ref[key]["count"]), key
    q = np.linspace(0.0, 0.5, 80)
    I_ref, P_ref = O.debye_ffgrid(ref, q, 0.25)
    # 1e-4 of I, or the bf16-split contraction's bound where the exv nearly
    # cancels the atoms (2e-8 of the profiles' absolute sum)
    err = np.abs(h.debye_transform(q) - I_ref)
    assert np.all(err <= 1e-4 * np.abs(I_ref) + 2e-8 * np.abs(P_ref).sum(0))
