"""Batched rigid-body scan (PartialHistogramManagerMT on many conformations)
against the oracle's partial manager, through the C ABI.

Bar: counts bit-exact per conformation; values within 1e-12 relative (FP64
sums of exact products); bin centres within the bound of axis_ok; I(q)
within 1e-5 relative of the FP64 oracle (stated bound 1e-4).
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O
from conftest import axis_ok

pytestmark = pytest.mark.gpu

I0 = {3: 6.0, 7: 7.0, 11: 8.0, 12: 16.0}  # C, N, O, S electron counts as XYZW weights


def make_bodies(sizes, seed, spacing=26.0):
    """Globules of the given sizes (radius ~ N^(1/3)) on a line, FP64, with I0 weights."""
    bodies, weights = [], []
    for k, n in enumerate(sizes):
        r = 20.0 * (n / 2000.0) ** (1 / 3)
        g = A.synth_globule(seed + k, n, (r, r, r), 2.0)
        xyz = g[:, :3].astype(np.float64) + np.array([k * spacing, 0.0, 0.0])
        bodies.append(xyz)
        weights.append(np.array([I0[int(t)] for t in g[:, 3].view(np.int32)], np.float32))
    return bodies, weights


def random_transforms(rng, bodies, n_conf, static=()):
    """Rotation about each body's centre + translation in [-10, 10]^3; bodies
    in `static` keep the identity in every conformation."""
    T = np.zeros((n_conf, len(bodies), 12))
    for c in range(n_conf):
        for b, xyz in enumerate(bodies):
            if b in static:
                T[c, b] = A.rigid_transform(np.eye(3), np.zeros(3))
                continue
            R = A.random_rotation(rng)
            cen = xyz.mean(0)
            t = cen - R @ cen + rng.uniform(-10, 10, 3)
            T[c, b] = A.rigid_transform(R, t)
    return T


def oracle_conf(bodies, weights, T, waters=None, bw=0.25):
    self_pts = np.concatenate([O.rigid_transform(b, w, A.rigid_transform(np.eye(3), np.zeros(3)))
                               for b, w in zip(bodies, weights)])
    moved = np.concatenate([O.rigid_transform(b, w, T[k]) for k, (b, w) in enumerate(zip(bodies, weights))])
    sizes = [len(b) for b in bodies]
    return O.hist_partial(self_pts, moved, sizes, waters, bw, 8000, nthreads=8)


def check_hist(h, ref, bw=0.25):
    assert h.max_bin == ref["max_bin"]
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT, "p_tot")):
        v, c, b = h.get(which)
        assert np.array_equal(c, ref[key]["count"]), (key, np.argwhere(c != ref[key]["count"])[:5])
        np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-12, err_msg=key)
        m = ref[key]["count"] > 0
        d, dr = b[m], ref[key]["bin_center"][m]
        assert np.all(np.abs(d - dr) <= 1e-5 * np.abs(dr) + 1e-6), key
    assert axis_ok(h.d_axis(), O.weighted_axis(ref["p_tot"], bw), bw, ref["p_tot"]["count"])


def relerr(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("static", [(), (0,)])
def test_rigid_scan_parity(static):
    # sizes with every residue mod 4: tail strips on every block
    bodies, weights = make_bodies([301, 498, 703], seed=11)
    rng = np.random.default_rng(5)
    wx = rng.normal(size=(150, 3)) * 15 + np.array([26.0, 0, 0])
    waters = np.zeros((150, 4), np.float32)
    waters[:, :3] = wx.astype(np.float32)
    waters[:, 3] = 1.0
    scan = A.RigidScan(bodies, weights, wx, np.ones(150, np.float32))
    T = random_transforms(rng, bodies, 5, static)
    q = np.linspace(0.01, 0.5, 60)
    I = scan.scan(T, q)
    for c in range(len(T)):
        ref = oracle_conf(bodies, weights, T[c], waters)
        check_hist(scan.histogram(c), ref)
        I_ref = O.debye_simple(ref, q, 0.25)
        err = relerr(I[c], I_ref)
        print(f"static {static} conformation {c}: max_bin {ref['max_bin']}, I(q) rel err {err:.2e}")
        assert err <= 1e-5


def test_rigid_scan_single_and_batch_agree():
    # a conformation scored alone (every block computed once) and inside a batch
    bodies, weights = make_bodies([250, 333, 417, 120], seed=21)
    rng = np.random.default_rng(8)
    T = random_transforms(rng, bodies, 6)
    q = np.linspace(0.0, 0.6, 40)
    scan = A.RigidScan(bodies, weights)
    I_batch = scan.scan(T, q)
    for c in (0, 3, 5):
        I_one = scan.scan(T[c:c + 1], q)
        assert np.array_equal(I_one[0], I_batch[c])
        ref = oracle_conf(bodies, weights, T[c])
        check_hist(scan.histogram(0), ref)


def test_rigid_graph_replay():
    # replays of the captured graph with new transforms between them
    bodies, weights = make_bodies([200, 260, 310], seed=31)
    rng = np.random.default_rng(9)
    T1 = random_transforms(rng, bodies, 4)
    T2 = random_transforms(rng, bodies, 4)
    q = np.linspace(0.01, 0.4, 30)
    scan = A.RigidScan(bodies, weights)
    a1 = scan.scan(T1, q)
    b1 = scan.scan(T2, q)
    a2 = scan.scan(T1, q)
    b2 = scan.scan(T2, q)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    assert not np.array_equal(a1, b1)
    for c in range(4):
        ref = oracle_conf(bodies, weights, T2[c])
        assert relerr(b2[c], O.debye_simple(ref, q, 0.25)) <= 1e-5


def test_rigid_autotune_matches_fixed_tiles():
    bodies, weights = make_bodies([400, 600], seed=41)
    rng = np.random.default_rng(10)
    T = random_transforms(rng, bodies, 8)
    q = np.linspace(0.01, 0.5, 50)
    fixed = A.RigidScan(bodies, weights).scan(T, q)
    tuned = A.RigidScan(bodies, weights, autotune=True)
    I = tuned.scan(T, q)
    info = tuned.info()
    print("autotuned:", info)
    assert info["tile_j"] in (1024, 2048, 4096, 8000)
    assert np.array_equal(I, fixed)


def test_rigid_scan_any_weights():
    # more than 64 distinct weights: the weighted engine's inter-body and
    # body-water blocks against the oracle's partial manager
    bodies, weights = make_bodies([301, 498, 257], seed=61)
    rng = np.random.default_rng(62)
    weights = [rng.uniform(0.5, 12.0, len(w)).astype(np.float32) for w in weights]
    wx = rng.normal(size=(90, 3)) * 15 + np.array([26.0, 0, 0])
    ww = rng.uniform(0.5, 1.5, 90).astype(np.float32)
    waters = np.zeros((90, 4), np.float32)
    waters[:, :3] = wx.astype(np.float32)
    waters[:, 3] = ww
    scan = A.RigidScan(bodies, weights, wx, ww)
    T = random_transforms(rng, bodies, 4, static=(1,))
    q = np.linspace(0.01, 0.5, 60)
    I = scan.scan(T, q)
    for c in range(len(T)):
        ref = oracle_conf(bodies, weights, T[c], waters)
        h = scan.histogram(c)
        assert h.max_bin == ref["max_bin"]
        for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT, "p_tot")):
            v, cnt, _ = h.get(which)
            assert np.array_equal(cnt, ref[key]["count"]), key
            np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-9, atol=1e-9 * np.max(ref[key]["value"]),
                                       err_msg=key)
        assert axis_ok(h.d_axis(), O.weighted_axis(ref["p_tot"], 0.25), 0.25, ref["p_tot"]["count"])
        err = relerr(I[c], O.debye_simple(ref, q, 0.25))
        print(f"conformation {c}: I(q) rel err {err:.2e}")
        assert err <= 1e-5


def test_rigid_overflow_is_an_error():
    bodies, weights = make_bodies([100, 100], seed=51)
    T = random_transforms(np.random.default_rng(1), bodies, 2)
    with pytest.raises(A.AusaxsError):
        A.RigidScan(bodies, weights, settings=A.default_settings(bin_count=20)).scan(T, [0.1])


def test_rigid_large_invariants():
    # the benchmark's bodies (3k/5k/8k/12k atoms) on a few conformations:
    # pair-count identities per conformation and agreement with single scans
    bodies, weights = make_bodies([3000, 5000, 8000, 12000], seed=4, spacing=45.0)
    rng = np.random.default_rng(4)
    T = random_transforms(rng, bodies, 3)
    scan = A.RigidScan(bodies, weights)
    q = A.default_q_axis()
    I = scan.scan(T, q)
    n = sum(len(b) for b in bodies)
    sizes = [len(b) for b in bodies]
    for c in range(3):
        h = scan.histogram(c)
        cnt = h.get(A.AA)[1].sum(dtype=np.int64)
        # every ordered pair once plus the per-body (unsigned) sum w^2 at bin 0
        self_terms = sum(int(np.sum(w.astype(np.float64) ** 2)) for w in weights)
        assert cnt == n * (n - 1) + self_terms
        assert np.all(np.isfinite(I[c]))
    assert sizes == [3000, 5000, 8000, 12000]


def test_rigid_benchmark_scan_512_conformations():
    # the benchmark's scan exactly: bodies of 3k/5k/8k/12k atoms, 512 random
    # conformations in one autotuned graph launch; pair-count identities on
    # every conformation, the oracle's partial manager on three of them, and a
    # replay of the captured graph bit-identical
    import bench

    cfg = bench.CONFIGS[4]
    bodies, weights, T = bench.rigid_inputs(cfg, 4)
    assert T.shape[0] == 512
    q = A.default_q_axis()
    scan = A.RigidScan(bodies, weights, settings=A.default_settings(bin_width=cfg["bw"]), autotune=True)
    I = scan.scan(T, q)
    info = scan.info()
    print("scan:", info)
    assert info["batch"] >= 512
    n = sum(len(b) for b in bodies)
    self_terms = sum(int(np.sum(w.astype(np.float64) ** 2)) for w in weights)
    for c in range(len(T)):
        assert scan.histogram(c).get(A.AA)[1].sum(dtype=np.int64) == n * (n - 1) + self_terms, c
    assert np.all(np.isfinite(I))
    for c in (0, 255, 511):
        ref = oracle_conf(bodies, weights, T[c], bw=cfg["bw"])
        check_hist(scan.histogram(c), ref, cfg["bw"])
        err = relerr(I[c], O.debye_simple(ref, q, cfg["bw"]))
        print(f"conformation {c}: I(q) rel err {err:.2e}")
        assert err <= 1e-5
    assert np.array_equal(scan.scan(T, q), I)


def water_shell(rng, centre, m, r0, r1):
    """m waters with radii uniform in [r0, r1] around centre (FP64 m x 3)."""
    v = rng.normal(size=(m, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return centre + v * rng.uniform(r0, r1, (m, 1))


def as_xyzw(xyz, w):
    out = np.zeros((len(xyz), 4), np.float32)
    out[:, :3] = np.asarray(xyz).astype(np.float32)
    out[:, 3] = w
    return out


@pytest.mark.parametrize("static", [(), (1,)])
def test_rigid_scan_hydrated_parity(static):
    # waters that change with every conformation (the re-hydrated step): aw of
    # every body and ww recomputed per conformation; water counts with every
    # residue mod 4, one conformation without waters
    bodies, weights = make_bodies([301, 498, 703], seed=71)
    rng = np.random.default_rng(72)
    w0 = water_shell(rng, np.array([26.0, 0, 0]), 150, 20, 30)
    scan = A.RigidScan(bodies, weights, w0, np.ones(150, np.float32))
    T = random_transforms(rng, bodies, 5, static)
    counts = [150, 97, 0, 203, 62]
    waters = [(water_shell(rng, np.array([26.0, 0, 0]) + rng.uniform(-3, 3, 3), m, 18, 32), 1.0) for m in counts]
    q = np.linspace(0.01, 0.5, 60)
    I = scan.scan_hydrated(T, waters, q)
    for c in range(len(T)):
        ref = oracle_conf(bodies, weights, T[c], as_xyzw(waters[c][0], 1.0))
        check_hist(scan.histogram(c), ref)
        err = relerr(I[c], O.debye_simple(ref, q, 0.25))
        print(f"static {static} conformation {c} ({counts[c]} waters): I(q) rel err {err:.2e}")
        assert err <= 1e-5


def test_rigid_scan_hydrated_same_waters_as_fixed():
    # the creation waters passed to every conformation give the fixed-water scan
    bodies, weights = make_bodies([250, 333, 417], seed=81)
    rng = np.random.default_rng(82)
    w0 = water_shell(rng, np.array([26.0, 0, 0]), 121, 20, 30)
    scan = A.RigidScan(bodies, weights, w0, np.ones(121, np.float32))
    T = random_transforms(rng, bodies, 4)
    q = np.linspace(0.0, 0.6, 50)
    I_fixed = scan.scan(T, q)
    h_fixed = [scan.histogram(c) for c in range(len(T))]
    I_hyd = scan.scan_hydrated(T, [(w0, 1.0)] * len(T), q)
    for c in range(len(T)):
        h = scan.histogram(c)
        for which in (A.AA, A.AW, A.WW, A.PTOT):
            a, b = h.get(which), h_fixed[c].get(which)
            assert np.array_equal(a[1], b[1])
            np.testing.assert_allclose(a[0], b[0], rtol=1e-14)
            np.testing.assert_allclose(a[2], b[2], rtol=1e-14)
    np.testing.assert_allclose(I_hyd, I_fixed, rtol=1e-13)
    # and the fixed-water scan is unchanged after a hydrated one
    assert np.array_equal(scan.scan(T, q), I_fixed)


def test_rigid_scan_hydrated_any_weights():
    # more than 64 distinct weights (weight pass) with per-conformation waters
    bodies, weights = make_bodies([301, 257], seed=91)
    rng = np.random.default_rng(92)
    weights = [rng.uniform(0.5, 12.0, len(w)).astype(np.float32) for w in weights]
    w0 = water_shell(rng, np.array([13.0, 0, 0]), 80, 18, 25)
    scan = A.RigidScan(bodies, weights, w0, rng.uniform(0.5, 1.5, 80).astype(np.float32))
    T = random_transforms(rng, bodies, 3)
    waters = []
    for m in (77, 0, 130):
        waters.append((water_shell(rng, np.array([13.0, 0, 0]), m, 17, 26), rng.uniform(0.5, 1.5, m).astype(np.float32)))
    q = np.linspace(0.01, 0.5, 60)
    I = scan.scan_hydrated(T, waters, q)
    for c in range(len(T)):
        ref = oracle_conf(bodies, weights, T[c], as_xyzw(*waters[c]))
        h = scan.histogram(c)
        assert h.max_bin == ref["max_bin"]
        for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT, "p_tot")):
            v, cnt, _ = h.get(which)
            assert np.array_equal(cnt, ref[key]["count"]), key
            np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-9, atol=1e-9 * np.max(np.abs(ref[key]["value"])),
                                       err_msg=key)
        assert axis_ok(h.d_axis(), O.weighted_axis(ref["p_tot"], 0.25), 0.25, ref["p_tot"]["count"])
        err = relerr(I[c], O.debye_simple(ref, q, 0.25))
        print(f"conformation {c}: I(q) rel err {err:.2e}")
        assert err <= 1e-5


def test_rigid_scan_hydrated_errors():
    bodies, weights = make_bodies([120, 130], seed=95)
    rng = np.random.default_rng(96)
    scan = A.RigidScan(bodies, weights)
    T = random_transforms(rng, bodies, 2)
    w = water_shell(rng, np.array([13.0, 0, 0]), 10, 15, 20)
    with pytest.raises(A.AusaxsError, match="water weight"):
        scan.scan_hydrated(T, [(w, 6.0), (w, 2.5)], [0.1])  # 2.5 is no weight of the scan
    with pytest.raises(A.AusaxsError, match="finite"):
        scan.scan_hydrated(T, [(w, 6.0), (np.full((3, 3), np.nan), 6.0)], [0.1])
    I = scan.scan_hydrated(T, [(w, 6.0), (w[:0], 6.0)], [0.1])
    assert np.all(np.isfinite(I))


def test_rigid_benchmark_bodies_hydrated():
    # the benchmark's bodies (3k/5k/8k/12k atoms) with a new hydration shell of
    # 3,000-4,000 waters in each of 64 conformations: pair-count identities on
    # every conformation, the oracle's partial manager on two
    import bench

    cfg = bench.CONFIGS[4]
    bodies, weights, T = bench.rigid_inputs(cfg, 4)
    T = T[:64]
    rng = np.random.default_rng(44)
    centre = np.concatenate(bodies).mean(0)
    w0 = water_shell(rng, centre, 3000, 40, 60)
    scan = A.RigidScan(bodies, weights, w0, np.ones(3000, np.float32),
                       settings=A.default_settings(bin_width=cfg["bw"]))
    counts = rng.integers(3000, 4001, len(T))
    waters = [(water_shell(rng, centre + rng.uniform(-5, 5, 3), int(m), 38, 62), 1.0) for m in counts]
    q = A.default_q_axis()
    I = scan.scan_hydrated(T, waters, q)
    print("scan:", scan.info())
    n = sum(len(b) for b in bodies)
    self_terms = sum(int(np.sum(w.astype(np.float64) ** 2)) for w in weights)
    for c in range(len(T)):
        h = scan.histogram(c)
        m = int(counts[c])
        assert h.get(A.AA)[1].sum(dtype=np.int64) == n * (n - 1) + self_terms, c
        assert h.get(A.AW)[1].sum(dtype=np.int64) == 2 * n * m, c
        assert h.get(A.WW)[1].sum(dtype=np.int64) == m * (m - 1) + m, c
    assert np.all(np.isfinite(I))
    for c in (0, 63):
        ref = oracle_conf(bodies, weights, T[c], as_xyzw(waters[c][0], 1.0), bw=cfg["bw"])
        check_hist(scan.histogram(c), ref, cfg["bw"])
        err = relerr(I[c], O.debye_simple(ref, q, cfg["bw"]))
        print(f"conformation {c}: I(q) rel err {err:.2e}")
        assert err <= 1e-5


def test_rigid_scan_hydrated_across_batches():
    # more conformations than one batch holds (1,024): the water sets of the
    # second batch follow those of the first in the input; each conformation
    # equals the same conformation scanned alone
    bodies, weights = make_bodies([60, 45], seed=97, spacing=14.0)
    rng = np.random.default_rng(98)
    scan = A.RigidScan(bodies, weights, np.zeros((1, 3)), np.ones(1, np.float32))
    n_conf = 1100
    T = random_transforms(rng, bodies, n_conf)
    counts = rng.integers(0, 24, n_conf)
    waters = [(water_shell(rng, np.array([7.0, 0, 0]), int(m), 12, 16), 1.0) for m in counts]
    q = np.linspace(0.01, 0.5, 20)
    I = scan.scan_hydrated(T, waters, q)
    assert np.all(np.isfinite(I))
    for c in (0, 1023, 1024, 1099):
        I1 = scan.scan_hydrated(T[c:c + 1], waters[c:c + 1], q)
        np.testing.assert_allclose(I1[0], I[c], rtol=1e-13)
    ref = oracle_conf(bodies, weights, T[1099], as_xyzw(waters[1099][0], 1.0))
    check_hist(scan.histogram(0), ref)
