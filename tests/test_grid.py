"""Grid excluded volume (host-side; no GPU): the product's grid builder and the
oracle's literal voxel-state restatement against an order-free closed form,
plus FFGrid bookkeeping invariants of the oracle."""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

VDW = {2: 1.1, 3: 1.96, 4: 1.96, 5: 1.96, 6: 1.96, 7: 1.79, 8: 1.79, 9: 1.79, 10: 1.79, 1: 1.71, 11: 1.71,
       12: 2.06, 13: 2.06, 14: 1.88}


def closed_form(atoms, ff, waters):
    """Dummies = atom-centre voxels u lattice points within max(r_vdw, 2.15) of an
    atom (FP64, ((dx^2 + dy^2) + dz^2)) that are not water-centre-only voxels;
    lexicographic order. Integer lattice: the grid origin is an integer."""
    centres = {tuple(int(v) for v in np.floor(a + 0.5)) for a in atoms}
    wcent = {tuple(int(v) for v in np.floor(w + 0.5)) for w in waters} - centres
    pts = set(centres)
    for a, f in zip(atoms, ff):
        r2 = max(VDW[int(f)] ** 2, 2.15 ** 2)
        c = np.floor(a + 0.5).astype(int)
        rng = range(-3, 4)
        for i in rng:
            x2 = (c[0] + i - a[0]) ** 2
            for j in rng:
                x2y2 = x2 + (c[1] + j - a[1]) ** 2
                for k in rng:
                    d = x2y2 + (c[2] + k - a[2]) ** 2
                    p = (c[0] + i, c[1] + j, c[2] + k)
                    if d <= r2 and p not in wcent:
                        pts.add(p)
    return np.array(sorted(pts), dtype=np.float64)


@pytest.mark.parametrize("seed,n,m", [(1, 60, 0), (2, 150, 40), (3, 300, 120)])
def test_grid_closed_form(seed, n, m):
    a = A.synth_globule(seed, n, (9, 9, 9), 2.0)
    xyz = a[:, :3].astype(np.float64)
    ff = a.view(np.int32)[:, 3]
    w = A.synth_shell(seed + 10, m, (9, 9, 9), 1.2)[:, :3].astype(np.float64) if m else np.zeros((0, 3))
    ref = closed_form(xyz, ff, w)
    d_orc, v_orc = O.grid_exv(xyz, ff, w)
    d_gpu, v_gpu = A.grid_exv(xyz, ff, w)
    assert np.array_equal(d_orc, ref)
    assert np.array_equal(d_gpu, ref)
    assert v_orc == v_gpu == len(ref)


def test_grid_water_blocking_by_hand():
    # one carbon at the origin; a water whose voxel is 2 A from the atom blocks
    # that voxel, one sharing the atom's own voxel does not
    atoms = np.array([[0.0, 0.0, 0.0]])
    ff = np.array([3])
    waters = np.array([[2.1, 0.0, 0.0], [0.2, 0.1, 0.0]])
    d, v = O.grid_exv(atoms, ff, waters)
    pts = {tuple(p) for p in d.astype(int)}
    assert (2, 0, 0) not in pts and (0, 0, 0) in pts and (1, 0, 0) in pts
    # ball of radius 2.15 around a lattice point has 33 lattice points (r^2 <= 4.6225: 0,1,2,3,4)
    assert len(pts) == 33 - 1
    d2, v2 = A.grid_exv(atoms, ff, waters)
    assert np.array_equal(d, d2) and v == v2


def test_ffgrid_oracle_invariants():
    a = A.synth_globule(4, 200, (9, 9, 9), 2.0)
    xyz = a[:, :3].astype(np.float64)
    ff = a.view(np.int32)[:, 3]
    w = A.synth_shell(5, 50, (9, 9, 9), 3.0)[:, :3].astype(np.float64)
    h = O.hist_ffgrid(xyz, ff, w)
    n, m, nx = len(xyz), len(w), len(h["dummies"])
    aa = h["aa"]
    # xx = [exv][exv]: n_x(n_x - 1) pair counts + n_x self
    assert aa["count"][0, 0].sum() == nx * (nx - 1) + nx
    # ax rows = [ff][exv]: every atom x dummy pair once
    assert aa["count"][1:, 0].sum() == n * nx
    # wx = aw[exv]
    assert h["aw"]["count"][0].sum() == m * nx
    # p_tot_ax = wx + ax
    assert h["p_tot_ax"]["count"].sum() == (n + m) * nx
    # FFGrid p_tot carries the FFAvg centres with zero counts
    assert np.all(h["p_tot"]["count"] == 0)
    assert h["volume"] == nx
