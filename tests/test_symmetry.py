"""Symmetry manager (SymmetryManagerMT, SymmetryManagerMT.cpp:39-170).

CPU tests pin the oracle: the reference's own analytic cases
(tests/feature/hist/symmetries/cyclic_symmetry.cpp, exact counts at known
distances with unit charges), pair schedules against closed forms and a brute
force over placement pairs, copy maps against closed-form rotations, and the
symmetry histogram against the same structure with every copy written out
(the reference's explicit_structure comparison). GPU tests compare the HIP
path with the oracle: counts bit-exact, values and bin centres to FP64
rounding, I(q) within 1e-4.
"""
import itertools
import math

import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

F32 = np.float32
R2, R3, R5, R8, R10, R11, R13, R18 = (math.sqrt(x) for x in (2, 3, 5, 8, 10, 11, 13, 18))
X, Y, Z = (1, 0, 0), (0, 1, 0), (0, 0, 1)


def cyc(offset, step, axis, angle, reps=1):
    return dict(offset=offset, step=step, axis=axis, angle=angle, reps=reps)


# (bodies, symmetries per body as cyc(...) dicts, waters, {distance: weighted count of p_tot})
# transcribed from tests/feature/hist/symmetries/cyclic_symmetry.cpp
CASES = {
    "one atom, no copies": ([[(0, 0, 0)]], [[]], [], {0: 1}),
    "one atom, one copy": ([[(0, 0, 0)]], [[cyc(0, X, Z, 0)]], [], {0: 2, 1: 2}),
    "one atom, two copies": ([[(0, 0, 0)]], [[cyc(0, X, Z, 0), cyc(0, (-1, 0, 0), Z, 0)]], [], {0: 3, 1: 4, 2: 2}),
    "one atom, four copies": ([[(0, 0, 0)]],
                              [[cyc(0, (-1, 0, 0), Z, 0), cyc(0, X, Z, 0), cyc(0, (0, -1, 0), Z, 0),
                                cyc(0, Y, Z, 0)]], [], {0: 5, 1: 8, R2: 8, 2: 4}),
    "two atoms, one copy": ([[(1, 0, 0), (0, 0, 0)]], [[cyc(0, Y, Z, 0)]], [], {0: 4, 1: 8, R2: 4}),
    "two atoms, two copies": ([[(1, 0, 0), (0, 0, 0)]], [[cyc(0, Y, Z, 0), cyc(0, (0, 2, 0), Z, 0)]], [],
                              {0: 6, 1: 14, R2: 8, 2: 4, R5: 4}),
    "two bodies, no copies": ([[(1, 0, 0)], [(0, 0, 0)]], [[], []], [], {0: 2, 1: 2}),
    "two bodies, copy of body 1": ([[(1, 0, 0)], [(0, 0, 0)]], [[cyc(0, Y, Z, 0)], []], [], {0: 3, 1: 4, R2: 2}),
    "two bodies, one copy each #1": ([[(1, 0, 0)], [(0, 0, 0)]], [[cyc(0, Y, Z, 0)], [cyc(0, Y, Z, 0)]], [],
                                     {0: 4, 1: 8, R2: 4}),
    "two bodies, one copy each #2": ([[(1, 0, 0)], [(0, 0, 0)]], [[cyc(0, X, Z, 0)], [cyc(0, (-1, 0, 0), Z, 0)]],
                                     [], {0: 4, 1: 6, 2: 4, 3: 2}),
    "waters, no copies": ([[(0, 0, 0)]], [[]], [(1, 0, 0)], {0: 2, 1: 2}),
    "waters, one copy": ([[(0, 0, 0)]], [[cyc(0, Y, Z, 0)]], [(1, 0, 0)], {0: 3, 1: 4, R2: 2}),
    "two repeats": ([[(0, 0, 0)]], [[cyc(0, X, Z, 0, 2)]], [], {0: 3, 1: 4, 2: 2}),
    "three repeats": ([[(0, 0, 0)]], [[cyc(0, X, Z, 0, 3)]], [], {0: 4, 1: 6, 2: 4, 3: 2}),
    "two bodies, two repeats": ([[(1, 0, 0)], [(0, 0, 0)]], [[cyc(0, Y, Z, 0, 2)], [cyc(0, Y, Z, 0, 2)]], [],
                                {0: 6, 1: 14, R2: 8, 2: 4, R5: 4}),
    "two bodies, different repeats": ([[(1, 0, 0)], [(0, 0, 0)]], [[cyc(0, Y, Z, 0, 1)], [cyc(0, Y, Z, 0, 2)]], [],
                                      {0: 5, 1: 10, R2: 6, 2: 2, R5: 2}),
    "rotation, one copy": ([[(1, 0, 0)]], [[cyc(X, 0, Y, math.pi / 2)]], [], {0: 2, R2: 2}),
    "rotation, three copies": ([[(1, 0, 0)]], [[cyc(X, 0, Y, math.pi / 2, 3)]], [], {0: 4, R2: 8, 2: 4}),
    "multi-atom cross": ([[(1, 0, 0), (2, 0, 0), (3, 0, 0)]], [[cyc((2, 0, 0), 0, Z, math.pi / 2, 3)]], [],
                         {0: 12, 1: 16, 2: 12, 3: 8, 4: 12, 5: 8, 6: 4, R2: 8, R5: 16, R8: 8, R10: 16, R13: 16,
                          R18: 8}),
    "helix": ([[(1, 0, 0)]], [[cyc(X, Z, Z, math.pi / 2, 4)]], [], {0: 5, R3: 8, R8: 6, R11: 4, 4: 2}),
}
# aa / aw / ww split of the water cases (same file, "one body with waters")
WATER_SPLIT = {
    "waters, no copies": dict(aa={0: 1}, aw={1: 2}, ww={0: 1}),
    "waters, one copy": dict(aa={0: 2, 1: 2}, aw={1: 2, R2: 2}, ww={0: 1}),
}


def vec3(v):
    return np.zeros(3) if np.isscalar(v) and v == 0 else np.asarray(v, np.float64)


def build(case):
    """Bodies (unit charges), Symmetry objects and waters of an analytic case."""
    bodies_xyz, syms, waters, _ = case
    bodies = [A.points_w(np.asarray(b, np.float64), 1.0) for b in bodies_xyz]
    sym_objs = []
    for b, per in zip(bodies_xyz, syms):
        cm = np.mean(np.asarray(b, np.float64), axis=0)
        sym_objs.append([A.Symmetry.cyclic(cm, vec3(s["offset"]), vec3(s["step"]), vec3(s["axis"]), s["angle"],
                                           s["reps"]) for s in per])
    w = A.points_w(np.asarray(waters, np.float64).reshape(-1, 3), 1.0) if waters else None
    return bodies, sym_objs, w


def expected_bins(spec, n, width=0.25):
    out = np.zeros(n)
    for d, c in spec.items():
        out[int(round(d / width))] += c
    return out


def oracle_of(bodies, syms, waters, **kw):
    return O.hist_symmetry(bodies, [[(s.transforms, s.schedule) for s in per] for per in syms], waters, **kw)


# the reference runs every analytic case through both managers
@pytest.mark.parametrize("partial", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reference_cases(name, partial):
    bodies, syms, waters = build(CASES[name])
    h = oracle_of(bodies, syms, waters, partial=partial)
    v = h["p_tot"]["value"]
    np.testing.assert_array_equal(v, expected_bins(CASES[name][3], len(v)))
    if name in WATER_SPLIT:
        for key, spec in WATER_SPLIT[name].items():
            np.testing.assert_array_equal(h[key]["value"], expected_bins(spec, len(v)))


def test_chain_schedule_closed_forms():
    # ISymmetry.cpp:43-66: every pair of the reps + 1 bodies is counted once
    for reps in range(1, 12):
        for closed in (False, True):
            s = A.chain_schedule(reps, closed)
            assert s[:, 2].sum() == reps * (reps + 1) // 2
            assert (s[:, 0] == 0).all() and (s[:, 1] >= 1).all() and (s[:, 1] <= reps).all()
    assert A.chain_schedule(1, True).tolist() == [[0, 1, 1]]
    assert A.chain_schedule(3, True).tolist() == [[0, 1, 4], [0, 2, 2]]
    assert A.chain_schedule(3, False).tolist() == [[0, 1, 3], [0, 2, 2], [0, 3, 1]]
    with pytest.raises(A.AusaxsError):
        A.chain_schedule(0, False)


def rot(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def tetrahedral_group():
    """The 12 proper rotations of the tetrahedron: cyclic permutation matrices
    times sign flips of even count."""
    out = []
    for perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        P = np.eye(3)[list(perm)]
        for signs in itertools.product((1, -1), repeat=3):
            if np.prod(signs) == 1:
                out.append(np.diag(signs) @ P)
    return out


def placements_of(rots, trans=None):
    trans = trans if trans is not None else [np.zeros(3)] * len(rots)
    return np.array([np.r_[R.ravel(), t] for R, t in zip(rots, trans)])


def test_placement_schedule_brute_force():
    rng = np.random.default_rng(3)
    body = rng.uniform(-3, 3, (7, 3)) + np.array([9.0, 4.0, 2.0])
    for rots in (tetrahedral_group(), [rot(Z, 2 * math.pi * k / 7) for k in range(7)]):
        n = len(rots)
        pl = placements_of(rots)
        s = A.placement_schedule(pl)
        assert s[:, 2].sum() == n * (n - 1) // 2
        # pairs of one class have the same inter-copy distances
        copies = [body @ R.T for R in rots]

        def dists(i, j):
            return np.sort(np.linalg.norm(copies[i][:, None] - copies[j][None], axis=2).ravel())
        classes = {}
        for i in range(n):
            for j in range(i + 1, n):
                key = tuple(np.round(dists(i, j), 6))
                classes[key] = classes.get(key, 0) + 1
        assert sorted(classes.values()) == sorted(s[:, 2].tolist())
        for a, b, _ in s:
            assert tuple(np.round(dists(a, b), 6)) in classes
    # cyclic group of 7: separations 1..3 each 7 times
    s = A.placement_schedule(placements_of([rot(Z, 2 * math.pi * k / 7) for k in range(7)]))
    assert sorted(s[:, 2].tolist()) == [7, 7, 7]


def test_cyclic_transform_closed_form():
    cm = np.array([2.0, -1.0, 0.5])
    t_i, t_r, axis, ang, reps = np.array([1.0, 2.0, 0.0]), np.array([0.0, 0.0, 0.7]), np.array([0, 0, 2.0]), 0.9, 4
    s = A.Symmetry.cyclic(cm, t_i, t_r, axis, ang, reps)
    R = rot(axis, ang)
    # copy 1: v -> R (v - cm) + R t_i - t_i + t_r + cm; copy k: copy 1 applied k times
    T1 = (cm - R @ cm) + (R @ t_i + t_r - t_i)
    v = np.array([0.3, 0.4, -1.2])
    w = v.copy()
    for k in range(reps):
        w = R @ w + T1
        T = s.transforms[k]
        got = T[:9].reshape(3, 3) @ v + T[9:]
        np.testing.assert_allclose(got, w, atol=1e-12)
    assert s.schedule[:, 2].sum() == reps * (reps + 1) // 2
    # closedness (CyclicSymmetry.cpp:109-118): pure rotations by 2 pi / (reps + 1)
    c5 = A.Symmetry.cyclic(cm, t_i, np.zeros(3), Z, 2 * math.pi / 5, 4)
    assert c5.schedule.tolist() == A.chain_schedule(4, True).tolist()
    open4 = A.Symmetry.cyclic(cm, t_i, np.zeros(3), Z, 2 * math.pi / 7, 4)
    assert open4.schedule.tolist() == A.chain_schedule(4, False).tolist()


def test_affine_copy_closed_forms():
    pts = np.array([[1.5, -2.25, 3.0, 2.0], [0.1, 0.2, 0.3, 1.0]], F32)
    ident = np.r_[np.eye(3).ravel(), [0, 0, 0]]
    np.testing.assert_array_equal(O.affine_float(pts, ident), pts)
    shift = np.r_[np.eye(3).ravel(), [1, -2, 0.5]]
    np.testing.assert_array_equal(O.affine_float(pts, shift)[:, :3],
                                  (pts[:, :3].astype(np.float64) + [1, -2, 0.5]).astype(F32))
    quarter = np.r_[np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], float).ravel(), [0, 0, 0]]
    q = O.affine_float(pts, quarter)
    np.testing.assert_array_equal(q[:, 0], -pts[:, 1])
    np.testing.assert_array_equal(q[:, 1], pts[:, 0])
    np.testing.assert_array_equal(q[:, 3], pts[:, 3])


def random_case(seed, n=(40, 25), weights=(1.0, 2.0, 0.5), n_waters=30):
    """Two bodies with weight classes: body 0 with a closed C5 rotation and an
    open translation chain, body 1 with a C3 rotation; waters around."""
    rng = np.random.default_rng(seed)
    b0 = rng.uniform(-4, 4, (n[0], 3)) + [12, 0, 0]
    b1 = rng.uniform(-3, 3, (n[1], 3)) + [0, 0, 25]
    bodies = [A.points_w(b, rng.choice(weights, len(b))) for b in (b0, b1)]
    syms = [[A.Symmetry.cyclic(b0.mean(0), np.zeros(3), np.zeros(3), Z, 2 * math.pi / 5, 4),
             A.Symmetry.cyclic(b0.mean(0), np.zeros(3), (0, 0, 9.0), Z, 0.3, 2)],
            [A.Symmetry.cyclic(b1.mean(0), (0, 6.0, 0), np.zeros(3), X, 2 * math.pi / 3, 2)]]
    waters = A.points_w(rng.uniform(-30, 30, (n_waters, 3)), 1.0) if n_waters else None
    return bodies, syms, waters


def test_oracle_matches_explicit_structure():
    """The symmetry histogram against the explicit structure through the plain
    manager (cyclic_symmetry.cpp, test_random): same pair multiset up to float
    rounding of the copies."""
    bodies, syms, waters = random_case(5, n_waters=0)
    h = oracle_of(bodies, syms, None, bin_width=0.25, bin_count=2000)
    expl = []
    for b, per in zip(bodies, syms):
        expl.append(b)
        for s in per:
            for T in s.transforms:
                expl.append(O.affine_float(b, T))
    ref = O.hist_simple(np.concatenate(expl), np.zeros((0, 4), F32), bin_width=0.25, bin_count=2000)
    n = min(len(h["p_tot"]), len(ref["p_tot"]))
    c0, c1 = h["p_tot"]["count"].astype(np.int64), ref["p_tot"]["count"].astype(np.int64)
    # bin 0 holds the truncated (unsigned) weight sums of the self terms,
    # per body in one case and over all copies in the other
    assert c0[1:].sum() == c1[1:].sum()
    # copy-pair classes differ by float rounding only: a few bin flips
    assert np.abs(c0[:n] - c1[:n]).sum() <= 0.01 * c0.sum()
    np.testing.assert_allclose(h["p_tot"]["value"].sum(), ref["p_tot"]["value"].sum(), rtol=1e-12)


def test_oracle_partial_matches_full():
    """PartialSymmetryManagerMT against SymmetryManagerMT on one state
    (partial_symmetry_manager.cpp compares them approximately): the same
    pairs in other row/column roles, so only the bins of tie pairs of the
    scalar-tail columns may differ."""
    bodies, syms, waters = random_case(9, n=(37, 22), n_waters=13)
    full = oracle_of(bodies, syms, waters)
    part = oracle_of(bodies, syms, waters, partial=True)
    c0, c1 = full["p_tot"]["count"].astype(np.int64), part["p_tot"]["count"].astype(np.int64)
    n = min(len(c0), len(c1))
    assert c0[1:].sum() == c1[1:].sum()
    assert np.abs(c0[:n] - c1[:n]).sum() <= 1e-3 * c0.sum()
    for key in ("aa", "aw", "ww"):
        assert full[key]["count"].astype(np.int64)[1:].sum() == part[key]["count"].astype(np.int64)[1:].sum()


def test_symmetry_input_errors():
    b = A.points_w(np.zeros((3, 3)), 1.0)
    T = np.r_[np.eye(3).ravel(), [1, 0, 0]]
    for sched in ([[0, 0, 1]], [[0, 2, 1]], [[0, 1, 0]], [[-1, 1, 1]]):
        with pytest.raises(A.AusaxsError):
            A.SymmetryStructure([b], [[A.Symmetry(T, sched)]])
    with pytest.raises(ValueError):
        A.SymmetryStructure([b], [])


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def compare(h, ref):
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww"), (A.PTOT, "p_tot")):
        v, c, b = h.get(which)
        assert np.array_equal(c, ref[key]["count"]), key
        np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-12, err_msg=key)
        m = ref[key]["count"] > 0
        np.testing.assert_allclose(b[m], ref[key]["bin_center"][m], rtol=1e-5, atol=1e-6, err_msg=key)


@pytest.mark.gpu
@pytest.mark.parametrize("partial", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_reference_cases(name, partial):
    bodies, syms, waters = build(CASES[name])
    h = A.SymmetryStructure(bodies, syms, waters, partial=partial).calculate()
    v, _, _ = h.get(A.PTOT)
    np.testing.assert_array_equal(v, expected_bins(CASES[name][3], len(v)))
    compare(h, oracle_of(bodies, syms, waters, partial=partial))


@pytest.mark.gpu
def test_gpu_partial_state_changes():
    """PartialSymmetryManagerMT across state changes (partial_symmetry_manager.cpp,
    test): a symmetry change, new waters, a rigid move of a body (its self term
    stays from the old coordinates), then an internal change (self term
    refreshed); each state against the oracle, counts bit-exact."""
    bodies, syms, waters = random_case(4, n=(300, 200), n_waters=120)
    st = A.SymmetryStructure(bodies, syms, waters, partial=True)
    compare(st.calculate(), oracle_of(bodies, syms, waters, partial=True, nthreads=4))
    self_bodies = [b.copy() for b in bodies]
    # modify a symmetry
    b0 = bodies[0]
    syms[0][1] = A.Symmetry.cyclic(b0[:, :3].mean(0), np.zeros(3), (0, 1.5, 9.0), Z, 0.5, 2)
    st.set_symmetry(0, 1, syms[0][1])
    compare(st.calculate(), oracle_of(bodies, syms, waters, partial=True, nthreads=4))
    # new waters
    waters = A.points_w(np.random.default_rng(6).uniform(-30, 30, (90, 3)), 1.0)
    st.set_waters(waters)
    compare(st.calculate(), oracle_of(bodies, syms, waters, partial=True, nthreads=4))
    # rigid move of body 1 (external): self term from the old coordinates
    moved = bodies[1].copy()
    moved[:, :3] = (moved[:, :3].astype(np.float64) @ rot(X, 0.4).T + [2.0, -1.0, 0.5]).astype(F32)
    bodies[1] = moved
    st.set_body(1, moved)
    compare(st.calculate(), oracle_of(bodies, syms, waters, partial=True, self_bodies=self_bodies, nthreads=4))
    # internal change of body 0 (weights): its self term is taken again
    b0 = bodies[0].copy()
    b0[:5, 3] = 3.0
    bodies[0] = b0
    self_bodies[0] = b0
    st.set_body(0, b0, internal=True)
    compare(st.calculate(), oracle_of(bodies, syms, waters, partial=True, self_bodies=self_bodies, nthreads=4))
    # the full manager recomputes the self term on any move
    full = A.SymmetryStructure(bodies, syms, waters)
    full.set_body(1, moved)
    compare(full.calculate(), oracle_of(bodies, syms, waters, nthreads=4))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_gpu_symmetry_parity(seed):
    bodies, syms, waters = random_case(seed, n=(700, 450), n_waters=300)
    st = A.SymmetryStructure(bodies, syms, waters)
    h = st.calculate()
    ref = oracle_of(bodies, syms, waters, nthreads=4)
    compare(h, ref)
    # calculate again on the resident copies: same result
    h2 = st.calculate()
    for w in (A.AA, A.AW, A.WW, A.PTOT):
        for x, y in zip(h.get(w), h2.get(w)):
            assert np.array_equal(x, y)
    q = np.linspace(0.0, 0.5, 101)
    I = h.debye_transform(q)
    I_ref = O.debye_simple(ref, q, 0.25)
    assert np.max(np.abs(I - I_ref) / np.abs(I_ref)) <= 1e-4


@pytest.mark.gpu
def test_gpu_placement_group():
    """A tetrahedral assembly (12 copies from placements, compute_pair_schedule)."""
    rng = np.random.default_rng(11)
    xyz = rng.uniform(-4, 4, (600, 3)) + [14, 5, 3]
    body = A.points_w(xyz, rng.choice([1.0, 6.0], len(xyz)))
    rots = tetrahedral_group()
    pl = placements_of(rots)
    sym = A.Symmetry(pl[1:], A.placement_schedule(pl))
    st = A.SymmetryStructure([body], [[sym]])
    h = st.calculate()
    ref = oracle_of([body], [[sym]], None, nthreads=4)
    compare(h, ref)
    # 12 copies of 600 atoms: 66 copy pairs in few classes, each evaluated once
    assert st.info()["pairs"] < 0.25 * (12 * 600) ** 2 / 2
