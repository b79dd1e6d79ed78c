"""GPU parity tests: the HIP path through the C ABI against the CPU oracle.

Bar: bin indices and counts bit-exact; summed distances (weighted bins)
within 1e-5 relative; I(q) within 1e-4 relative of the FP64 oracle.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O
from conftest import axis_ok, compare_hist_approx, cube_atoms, globule

pytestmark = pytest.mark.gpu

F32 = np.float32


def rel_close(a, b, rtol, atol=1e-9):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= atol + rtol * np.abs(b))


def test_device_is_gfx950():
    info = A.device_info()
    print(info)
    assert info["arch"].startswith("gfx950")


@pytest.mark.parametrize("bw", [0.25, 0.1, 1.0, 0.05, 0.3, 2.0])
def test_binning_arithmetic_exhaustive(bw):
    # every float32 d2 up to the widest histogram the host accepts (2^15 bins,
    # distances to 32768 bin widths): the fast path (rsq root, window test,
    # correctly rounded fallback near ties) bins like the reference's
    # sqrt_rn rule, the device sqrtf is correctly rounded (verified in FP64),
    # the fallback's correction equals it on every normal d2, and every offset
    # code is within one unit of the real root's offset
    hi = int(np.array((32768 * bw) ** 2, np.float32).view(np.uint32))
    bad_bin, bad_cr, bad_fix, bad_off, _, _ = A.bin_check(0, hi, bw)
    assert (bad_bin, bad_cr, bad_fix, bad_off) == (0, 0, 0, 0)


@pytest.mark.parametrize("bw", [0.25, 0.1, 1.0])
def test_offsets_unbiased_at_long_distances(bw):
    # every float d2 in [1.6e4, 6.4e4) (distances 126-253 A, where one ulp is
    # 2^-16 A and v_sqrt_f32's shortfall alone would move a bin centre by up to
    # ~5e-6 A): the pair kernel's offset codes average to the real root's
    # offsets within 0.01 units, as do the correctly rounded roots the
    # reference sums, and no code is more than one unit off
    lo = int(np.array(1.6e4, np.float32).view(np.uint32))
    hi = int(np.array(6.4e4, np.float32).view(np.uint32)) - 1
    bad_bin, _, _, bad_off, mean_err, mean_err_cr = A.bin_check(lo, hi, bw)
    assert bad_bin == 0 and bad_off == 0
    assert abs(mean_err) < 0.01, mean_err
    assert abs(mean_err_cr) < 0.01, mean_err_cr


def check_pairs(a, b, bw, nb):
    cnt, ds = A.pair_histogram(a, b, bw, nb)
    ref = O.pair_hist(a, b, bw, nb, factor=1)
    assert np.array_equal(cnt, ref["count"].astype(np.uint64))
    m = ref["count"] > 0
    assert rel_close(ds[m], ref["bin_center"][m], 1e-5, atol=1e-6)
    assert np.all(ds[~m] == 0)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5, 7, 64, 100, 513, 1027, 2049])
def test_self_pairs_random(n):
    rng = np.random.default_rng(n)
    a = np.zeros((n, 4), F32)
    a[:, :3] = rng.standard_normal((n, 3)) * 12
    check_pairs(a, None, 0.25, 800)


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (100, 37), (513, 514), (700, 1029), (1500, 4)])
def test_cross_pairs_random(n, m):
    rng = np.random.default_rng(n * 7 + m)
    a = np.zeros((n, 4), F32)
    b = np.zeros((m, 4), F32)
    a[:, :3] = rng.standard_normal((n, 3)) * 12
    b[:, :3] = rng.standard_normal((m, 3)) * 12 + 3
    check_pairs(a, b, 0.1, 1200)


@pytest.mark.parametrize("n", [9, 10, 11, 12, 131])
def test_exact_ties_follow_reference_rules(n):
    # points 0.625 A apart on a line: x = 2.5 k, an exact tie for odd k, where the
    # SIMD body (half to even) and the scalar tail (half up) disagree
    a = np.zeros((n, 4), F32)
    a[:, 0] = np.arange(n) * F32(0.625)
    check_pairs(a, None, 0.25, 400)
    check_pairs(a, a[::-1].copy(), 0.25, 400)


def test_integer_lattice():
    g = np.stack(np.meshgrid(np.arange(12), np.arange(9), np.arange(7), indexing="ij"), -1).reshape(-1, 3)
    a = np.zeros((len(g), 4), F32)
    a[:, :3] = g
    check_pairs(a, None, 0.25, 200)
    check_pairs(a, None, 1.0, 60)


# --------------------------------------------------------------------------
# managers
# --------------------------------------------------------------------------
def compare_entries(gpu, ref):
    v, c, b = gpu
    assert np.array_equal(c, ref["count"]), np.argwhere(c != ref["count"])[:5]
    assert np.array_equal(v, ref["value"])
    m = ref["count"] > 0
    assert rel_close(b[m], ref["bin_center"][m], 1e-5, atol=1e-6)


def compare_ff_hist(h, ref, mode):
    assert h.max_bin == ref["max_bin"]
    if mode == A.MODE_SIMPLE:
        for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT, "p_tot")):
            v, c, b = h.get(which)
            assert np.array_equal(c, ref[key]["count"]), key
            np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-12)
            m = ref[key]["count"] > 0
            assert rel_close(b[m], ref[key]["bin_center"][m], 1e-5, atol=1e-6)
        return
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww")):
        v, c, b = h.get(which)
        rv = ref[key]
        assert np.array_equal(c, rv["count"]), key
        np.testing.assert_allclose(v, rv["value"], rtol=1e-14, err_msg=key)
        m = rv["count"] > 0
        assert rel_close(b[m], rv["bin_center"][m], 1e-5, atol=1e-6), key
    v, c, b = h.get(A.PTOT)
    assert np.array_equal(c, ref["p_tot"]["count"])
    m = ref["p_tot"]["count"] > 0
    assert rel_close(b[m], ref["p_tot"]["bin_center"][m], 1e-5, atol=1e-6)
    d = h.d_axis()
    np.testing.assert_allclose(d, O.weighted_axis(ref["p_tot"], 0.25 if len(d) else 0.25), rtol=1e-5, atol=1e-6)


def config1(n=2000, waters=300):
    atoms = globule(n, 20.0, 1)
    w = A.synth_shell(5, waters, (20, 20, 20), 3.0)
    return atoms, w


@pytest.mark.parametrize("mode", [A.MODE_FFEXPLICIT, A.MODE_FFAVG])
def test_ff_manager_parity(mode):
    atoms, waters = config1()
    z = 0.37
    h = A.calculate_all(mode, atoms, waters, z_exv=z)
    ref = O.hist_ff(O.FFAVG if mode == A.MODE_FFAVG else O.FFEXPLICIT, atoms, waters, z_exv=z, nthreads=4)
    compare_ff_hist(h, ref, mode)


def test_ffavg_excluded_volume_charge_from_grid():
    # HistogramManagerMTFFAvg.cpp:355-359: Z_exv = V_grid * 0.334 / N with the
    # molecule's grid volume, here from FP64 atoms and waters; the oracle's
    # grid gives the same volume, and the histograms and I(q) follow
    a = globule(1500, 18.0, 12)
    xyz, ff = a[:, :3].astype(np.float64), a[:, 3].view(np.int32).copy()
    w = A.synth_shell(13, 400, (18, 18, 18), 3.0)[:, :3].astype(np.float64)
    st = A.FFAvgStructure(xyz, ff, w)
    vol = O.grid_exv(xyz, ff, w)[1]
    z = vol * 0.334 / len(xyz)
    assert st.z_exv == z, (st.z_exv, z)
    h = st.calculate()
    ref = O.hist_ff(O.FFAVG, O.pack_ff(xyz, ff), O.pack_ff(w, np.ones(len(w), np.int32)), z_exv=z, nthreads=4)
    compare_ff_hist(h, ref, A.MODE_FFAVG)
    q = np.linspace(0.0, 0.5, 100)
    I_ref, _ = O.debye_ff(O.FFAVG, ref, q, 0.25)
    assert relerr(h.debye_transform(q), I_ref) <= 1e-4


def test_ff_manager_types_and_tails():
    # every active atom type, sizes with all residues mod 4, types interleaved
    rng = np.random.default_rng(4)
    for n, m in ((37, 0), (38, 5), (41, 6), (400, 3)):
        xyz = rng.standard_normal((n, 3)) * 9
        ff = rng.choice([2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14], n)
        atoms = A.points_ff(xyz, ff)
        waters = A.points_ff(rng.standard_normal((m, 3)) * 11, 1)
        for mode in (A.MODE_FFEXPLICIT, A.MODE_FFAVG):
            h = A.calculate_all(mode, atoms, waters, z_exv=1.3)
            ref = O.hist_ff(O.FFAVG if mode == A.MODE_FFAVG else O.FFEXPLICIT, atoms, waters, z_exv=1.3)
            compare_ff_hist(h, ref, mode)


def test_simple_manager_parity():
    rng = np.random.default_rng(8)
    atoms = globule(1200, 16.0, 3)
    atoms[:, 3] = rng.choice([6.0, 7.0, 8.0, 5.5], len(atoms)).astype(F32)
    waters = A.synth_shell(6, 203, (16, 16, 16), 3.0)
    waters[:, 3] = 1.0
    h = A.calculate_all(A.MODE_SIMPLE, atoms, waters)
    ref = O.hist_simple(atoms, waters)
    compare_ff_hist(h, ref, A.MODE_SIMPLE)


@pytest.mark.parametrize("n,m,lo,hi", [(1200, 203, 0.5, 16.0), (777, 0, -3.0, 9.0), (3001, 1499, 0.0, 1.0)])
def test_simple_manager_any_weights(n, m, lo, hi):
    # one weight per point (occupancies, effective charges: more than 64
    # distinct values, zeros and negative ones included), as
    # HistogramManagerMT / SimpleCalculator add<2> take them
    # (SimpleCalculator.h:195-207): the weighted engine's values are the
    # oracle's FP64 sums of float products to 1e-9
    rng = np.random.default_rng(n + m)
    atoms = globule(n, 16.0 * (n / 1200) ** (1 / 3), 9)
    atoms[:, 3] = rng.uniform(lo, hi, n).astype(F32)
    atoms[::17, 3] = 0.0
    waters = A.synth_shell(7, max(m, 1), (16, 16, 16), 3.0)[:m]
    waters[:, 3] = rng.uniform(0.2, 1.0, m).astype(F32)
    h = A.calculate_all(A.MODE_SIMPLE, atoms, waters)
    ref = O.hist_simple(atoms, waters, nthreads=4)
    assert h.max_bin == ref["max_bin"]
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT, "p_tot")):
        v, c, b = h.get(which)
        rv = ref[key]
        assert np.array_equal(c, rv["count"]), key
        scale = np.max(np.abs(rv["value"])) or 1.0
        np.testing.assert_allclose(v, rv["value"], rtol=1e-9, atol=1e-12 * scale, err_msg=key)
    assert axis_ok(h.d_axis(), O.weighted_axis(ref["p_tot"], 0.25), 0.25, ref["p_tot"]["count"])
    q = np.linspace(0.0, 0.5, 100)
    I = h.debye_transform(q)
    I_ref = O.debye_simple(ref, q, 0.25)
    err = np.max(np.abs(I - I_ref)) / np.max(np.abs(I_ref))
    print(f"weights in [{lo}, {hi}]: I(q) max error {err:.2e} of max |I|")
    assert err <= 1e-6


def test_cube_known_counts():
    atoms = A.points_ff(cube_atoms(), 3)
    waters = A.points_ff(cube_atoms(), 1)
    h = A.calculate_all(A.MODE_FFEXPLICIT, atoms, waters)
    v, c, b = h.get(A.AA)
    assert h.max_bin == 15
    assert list(v[3, 3, [0, 8, 11, 14]]) == [8, 24, 24, 8]
    assert v.sum() == 64
    d = h.d_axis()
    assert abs(d[8] - 2) < 1e-6 and abs(d[11] - 8 ** 0.5) < 1e-6 and abs(d[14] - 12 ** 0.5) < 1e-6


def test_determinism_across_tilings():
    atoms, waters = config1(3000, 500)
    outs = []
    for tj in (256, 1000, 8064):
        h = A.calculate_all(A.MODE_FFEXPLICIT, atoms, waters, settings=A.default_settings(tile_j=tj))
        outs.append([h.get(w) for w in (A.AA, A.AW, A.WW)])
    for o in outs[1:]:
        for (v0, c0, b0), (v1, c1, b1) in zip(outs[0], o):
            assert np.array_equal(c0, c1) and np.array_equal(v0, v1) and np.array_equal(b0, b1)


def test_bin_overflow_is_an_error():
    a = A.points_ff(np.array([[0, 0, 0], [30, 0, 0]]), 3)
    with pytest.raises(A.AusaxsError, match="bin_count"):
        A.calculate_all(A.MODE_FFEXPLICIT, a, settings=A.default_settings(bin_count=100))


def test_unknown_form_factor_is_an_error():
    a = A.points_ff(np.zeros((3, 3)), 0)
    with pytest.raises(A.AusaxsError, match="form factor"):
        A.calculate_all(A.MODE_FFEXPLICIT, a)


# --------------------------------------------------------------------------
# Debye
# --------------------------------------------------------------------------
def relerr(x, ref):
    return np.max(np.abs(x - ref) / np.abs(ref))


@pytest.mark.parametrize("mode", [A.MODE_FFEXPLICIT, A.MODE_FFAVG])
def test_intensity_parity(mode):
    atoms, waters = config1()
    omode = O.FFAVG if mode == A.MODE_FFAVG else O.FFEXPLICIT
    ref_h = O.hist_ff(omode, atoms, waters, z_exv=0.5, nthreads=4)
    q = np.linspace(0.0, 0.5, 200)
    params = dict(cw=1.1, cx=1.05, crho=0.97, dw_atomic=0.2, dw_exv=0.1)
    I_ref, P_ref = O.debye_ff(omode, ref_h, q, 0.25, long_double=True, **params)
    # terms 0 (FP64 on device) differs from the oracle only through the bin
    # centres (offsets of the real root in units of 2^-22 A against sums of
    # correctly rounded roots), which I(q) amplifies near its minima (it spans
    # ~4 decades here)
    for terms, tol in ((3, 1e-4), (0, 2e-5)):
        h = A.calculate_all(mode, atoms, waters, settings=A.default_settings(debye_terms=terms), z_exv=0.5)
        assert axis_ok(h.d_axis(), O.weighted_axis(ref_h["p_tot"], 0.25), 0.25, ref_h["p_tot"]["count"])
        h.set_params(**params)
        I, P = h.debye_transform(q, profiles=True)
        err = relerr(I, I_ref)
        print(f"mode {mode} terms {terms}: max rel err {err:.3e}")
        assert err <= tol


def test_debye_contract_wide_angle():
    # 250k-atom-like histogram shape, WAXS q up to 1.5, 0.1 A bins
    atoms = globule(6000, 30.0, 2, 1.8)
    ref_h = O.hist_ff(O.FFEXPLICIT, atoms, None, bin_width=0.1, nthreads=8)
    mb = ref_h["max_bin"]
    d = O.weighted_axis(ref_h["p_tot"], 0.1)
    H = ref_h["aa"]["value"].reshape(-1, mb)
    H = H[np.any(H != 0, axis=1)]
    q = np.linspace(0.0015, 1.5, 1000)
    ref = O.contract(H, O.sinc_table(q, d), long_double=True)
    I_ref = ref.sum(0)
    for terms, tol in ((3, 1e-4), (2, None)):
        out = A.debye_contract(H, d, q, terms)
        err = relerr(out.sum(0), I_ref)
        print(f"terms {terms}: I(q) max rel err {err:.3e}")
        if tol:
            assert err <= tol
    f64 = A.debye_contract(H, d, q, 0)
    assert relerr(f64.sum(0), I_ref) <= 1e-10


def test_golden_carbon_sphere_gpu(golden_carbon):
    xyz, q, I_ref = golden_carbon
    atoms = A.points_w(xyz, 1.0)
    h = A.calculate_all(A.MODE_SIMPLE, atoms, None, settings=A.default_settings(bin_width=0.1))
    I = h.debye_transform(q) * np.exp(q * q)
    ok, i = compare_hist_approx(I, I_ref, 1e-6, 0.01)
    assert ok, (i, I[i], I_ref[i])


def test_large_structure_invariants():
    # beyond oracle sizes: exact pair-count identities
    atoms = A.synth_globule(3, 60000, (60, 45, 35), 1.8)
    waters = A.synth_shell(4, 9000, (60, 45, 35), 3.0)
    h = A.calculate_all(A.MODE_FFEXPLICIT, atoms, waters, settings=A.default_settings(bin_width=0.1))
    n, m = len(atoms), len(waters)
    assert h.get(A.AA)[1].sum(dtype=np.int64) == n * (n - 1) + n
    assert h.get(A.AW)[1].sum(dtype=np.int64) == n * m
    assert h.get(A.WW)[1].sum(dtype=np.int64) == m * (m - 1) + m


# --------------------------------------------------------------------------
# grid excluded volume (HistogramManagerMTFFGrid)
# --------------------------------------------------------------------------
def test_ffgrid_parity():
    a = A.synth_globule(6, 700, (12, 12, 12), 2.0)
    xyz = a[:, :3].astype(np.float64)
    ff = a.view(np.int32)[:, 3]
    w = A.synth_shell(7, 180, (12, 12, 12), 3.0)[:, :3].astype(np.float64)
    h = A.GridStructure(xyz, ff, w).calculate()
    ref = O.hist_ffgrid(xyz, ff, w, nthreads=4)
    assert h.max_bin == ref["max_bin"]
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT_AX, "p_tot_ax"), (A.XX, "p_xx")):
        v, c, b = h.get(which)
        rv = ref[key]
        assert np.array_equal(c, rv["count"]), key
        np.testing.assert_allclose(v, rv["value"], rtol=1e-14, err_msg=key)
        m = rv["count"] > 0
        assert rel_close(b[m], rv["bin_center"][m], 1e-5, atol=1e-6), key
    v, c, b = h.get(A.PTOT)
    np.testing.assert_allclose(v, ref["p_tot"]["value"], rtol=1e-12)
    assert np.all(c == 0)
    np.testing.assert_allclose(b, ref["p_tot"]["bin_center"], rtol=1e-6, atol=1e-7)
    d_ax, d_xx = h.grid_axes()
    assert axis_ok(d_ax, ref["d_ax"], 0.25) and axis_ok(d_xx, ref["d_xx"], 0.25)
    q = np.linspace(0.0, 0.5, 150)
    params = dict(cw=1.05, cx=0.98, crho=1.02, dw_atomic=0.1, dw_exv=0.05)
    I_ref, _ = O.debye_ffgrid(ref, q, 0.25, long_double=True, **params)
    h.set_params(**params)
    err = relerr(h.debye_transform(q), I_ref)
    print(f"ffgrid I(q) max rel err {err:.3e}")
    assert err <= 1e-4


def test_ffgrid_large_invariants():
    # 20k-atom globule (radius 45 A) + 6k waters: exact pair-count identities
    a = A.synth_globule(2, 20000, (45, 45, 45), 2.0)
    xyz = a[:, :3].astype(np.float64)
    ff = a.view(np.int32)[:, 3]
    w = A.synth_shell(3, 6000, (45, 45, 45), 3.0)[:, :3].astype(np.float64)
    nx = len(A.grid_exv(xyz, ff, w)[0])
    h = A.GridStructure(xyz, ff, w).calculate()
    n, m = len(xyz), len(w)
    _, c, _ = h.get(A.AA)
    assert c[0, 0].sum(dtype=np.int64) == nx * (nx - 1) + nx
    assert c[1:, 0].sum(dtype=np.int64) == n * nx
    assert h.get(A.AW)[1][0].sum(dtype=np.int64) == m * nx


def test_cpp_consumer_matches_python_api():
    # the C++ caller of the header (tests/capi_consumer.cpp, built by the
    # package build) copies the FF-explicit and FF-avg distributions into the
    # reference's layouts and runs I(q); its figures equal those of the same
    # structures through the Python binding, sums taken in the same order
    import json
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent / "_bin" / "capi_consumer"
    assert exe.exists(), f"{exe} is not built; run __graft_entry__.build()"
    n = 1500
    print("capi_consumer", n, flush=True)
    r = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout)
    m = n // 4
    atoms = A.synth_globule(7, n, (18, 16, 14), 2.0)
    waters = A.synth_shell(8, m, (18, 16, 14), 3.0)
    q = A.default_q_axis(1e-4, 0.5)
    assert out["nq"] == len(q)

    def seqsum(x):
        return float(np.cumsum(np.asarray(x, np.float64).ravel())[-1])

    def check(key, h):
        o = out[key]
        assert o["max_bin"] == h.max_bin
        assert o["aa_sum"] == seqsum(h.get(A.AA)[0])
        assert o["aw_sum"] == seqsum(h.get(A.AW)[0])
        assert o["ww_sum"] == seqsum(h.get(A.WW)[0])
        v, c, b = h.get(A.PTOT)
        assert o["ptot_value_sum"] == seqsum(v)
        assert o["ptot_count_sum"] == int(c.astype(np.int64).sum())
        assert o["ptot_center_sum"] == seqsum(b)
        assert o["d_sum"] == seqsum(h.d_axis())
        assert np.array_equal(np.array(o["I"]), h.debye_transform(q))

    check("ffexplicit", A.calculate_all(A.MODE_FFEXPLICIT, atoms, waters, z_exv=0.0))
    xyz = atoms[:, :3].astype(np.float64)
    st = A.FFAvgStructure(xyz, atoms[:, 3].view(np.int32).copy(), waters[:, :3].astype(np.float64))
    assert out["z_exv"] == st.z_exv
    check("ffavg", st.calculate())
