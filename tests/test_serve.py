"""Serving loop: continuous batching of mixed profile requests.

Each request's I(q) must be the same alone and in a mix of concurrent
requests (bit-identical: integer histograms and deterministic kernels), equal
to the direct structure API on the same input, and within 1e-4 relative of
the FP64 oracle.
"""
import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

MODES = (A.MODE_FFAVG, A.MODE_FFEXPLICIT, A.MODE_FFGRID)


def make_request(seed, n, mode):
    r = 20.0 * (n / 2000.0) ** (1 / 3)
    g = A.synth_globule(seed, n, (r, r, r), 2.0)
    w = A.synth_shell(seed + 1000, max(1, n // 4), (r, r, r), 3.0)
    return dict(mode=mode, xyz=g[:, :3].astype(np.float64), ff=g[:, 3].view(np.int32).copy(),
                wxyz=w[:, :3].astype(np.float64), q=np.linspace(0.0, 0.5, 200))


def direct(req):
    """The same evaluation through the structure API (no server)."""
    if req["mode"] == A.MODE_FFGRID:
        st = A.GridStructure(req["xyz"], req["ff"], req["wxyz"])
    else:
        atoms = A.points_ff(req["xyz"], req["ff"])
        waters = A.points_ff(req["wxyz"], 1)
        st = A.Structure(req["mode"], atoms, waters, z_exv=1.0)
    return st.calculate().debye_transform(req["q"])


def requests(count, seed=1):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(200), np.log(3000), count)).astype(int)
    return [make_request(seed * 100 + k, int(n), MODES[k % 3]) for k, n in enumerate(sizes)]


def test_serve_same_alone_and_in_mix():
    reqs = requests(15)
    srv = A.Server(slots=4, budget_gb=8.0)
    ids = [srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]) for r in reqs]
    mixed = [srv.wait(i) for i in ids]
    st = srv.stats()
    print("mix stats:", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    alone_srv = A.Server(slots=1, budget_gb=8.0)
    for r, res in zip(reqs, mixed):
        I_mix = res[0]
        I_alone = alone_srv.wait(alone_srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))[0]
        assert np.array_equal(I_mix, I_alone), r["mode"]
        assert np.array_equal(I_mix, direct(r)), r["mode"]


def test_serve_against_oracle():
    srv = A.Server(slots=3, budget_gb=4.0)
    reqs = [make_request(7, 400, A.MODE_FFAVG), make_request(8, 350, A.MODE_FFEXPLICIT),
            make_request(9, 300, A.MODE_FFGRID)]
    ids = [srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]) for r in reqs]
    for r, i in zip(reqs, ids):
        I = srv.wait(i)[0]
        if r["mode"] == A.MODE_FFGRID:
            h = O.hist_ffgrid(r["xyz"], r["ff"], r["wxyz"], 0.25, 8000, nthreads=4)
            I_ref, _ = O.debye_ffgrid(h, r["q"], 0.25)
        else:
            om = O.FFAVG if r["mode"] == A.MODE_FFAVG else O.FFEXPLICIT
            atoms = O.pack_ff(r["xyz"], r["ff"])
            waters = O.pack_ff(r["wxyz"], np.ones(len(r["wxyz"]), np.int32))
            h = O.hist_ff(om, atoms, waters, 0.25, 8000, 1.0, 4)
            I_ref, _ = O.debye_ff(om, h, r["q"], 0.25)
        err = float(np.max(np.abs(I - I_ref) / np.abs(I_ref)))
        print(f"mode {r['mode']}: rel err {err:.2e}")
        assert err <= 1e-4


def test_serve_admission_by_memory():
    reqs = requests(8, seed=3)
    # a budget a little above the largest request: at most two fit at once
    sizes = []
    for r in reqs:
        srv1 = A.Server(slots=1, budget_gb=64.0)
        srv1.wait(srv1.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))
        sizes.append(srv1.stats()["peak_gb"])
        srv1.close()
    budget = 1.5 * max(sizes)
    srv = A.Server(slots=8, budget_gb=budget)
    ids = [srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]) for r in reqs]
    for i in ids:
        assert srv.wait(i) is not None
    st = srv.stats()
    print("admission stats:", st, "budget GiB", budget)
    assert st["completed"] == len(reqs)
    assert st["peak_gb"] <= budget + 1e-9
    # a request larger than the whole budget is refused
    tiny = A.Server(slots=2, budget_gb=1e-4)
    big = reqs[int(np.argmax(sizes))]
    with pytest.raises(A.AusaxsError):
        tiny.submit(big["mode"], big["xyz"], big["ff"], big["wxyz"], big["q"])


def oracle_intensity(mode, xyz, ff, wxyz, q, nthreads=8):
    if mode == A.MODE_FFGRID:
        h = O.hist_ffgrid(xyz, ff, wxyz, 0.25, 8000, nthreads=nthreads)
        return O.debye_ffgrid(h, q, 0.25)[0]
    om = O.FFAVG if mode == A.MODE_FFAVG else O.FFEXPLICIT
    h = O.hist_ff(om, O.pack_ff(xyz, ff), O.pack_ff(wxyz, np.ones(len(wxyz), np.int32)), 0.25, 8000, 1.0, nthreads)
    return O.debye_ff(om, h, q, 0.25)[0]


def test_serve_benchmark_requests():
    # the benchmark's request stream: structures of its pool (log-uniform
    # 1k-100k atoms, 30 % waters), modes uniform over FF-avg / FF-explicit /
    # FF-grid, 200 q, Poisson arrivals while earlier requests run; each result
    # equals the same request served alone, and the oracle's for requests up to
    # 20k atoms (6k for grid requests); the larger ones are finite (the
    # contrast profiles' sum may change sign in the tail)
    import time

    import bench

    cfg = bench.CONFIGS[5]
    pool = bench.serve_pool(cfg, 5)
    rng = np.random.default_rng(6)
    q = np.linspace(0.0, 0.5, 200)
    reqs = [(int(rng.integers(len(pool))), MODES[int(rng.integers(3))]) for _ in range(40)]
    gaps = rng.exponential(0.01, len(reqs))
    srv = A.Server(slots=cfg["slots"], budget_gb=cfg["budget_gb"])
    ids = []
    for (k, mode), g in zip(reqs, gaps):
        xyz, ff, w = pool[k]
        ids.append(srv.submit(mode, xyz, ff, w, q))
        time.sleep(g)
    results = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    print("stats:", st)
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    alone = A.Server(slots=1, budget_gb=cfg["budget_gb"])
    checked = 0
    for (k, mode), I in zip(reqs, results):
        xyz, ff, w = pool[k]
        assert np.array_equal(I, alone.wait(alone.submit(mode, xyz, ff, w, q))[0]), (k, mode)
        assert np.all(np.isfinite(I))
        if len(xyz) <= (6000 if mode == A.MODE_FFGRID else 20000):
            I_ref = oracle_intensity(mode, xyz, ff, w, q)
            err = float(np.max(np.abs(I - I_ref) / np.abs(I_ref)))
            print(f"request of {len(xyz)} atoms, mode {mode}: rel err {err:.2e}")
            assert err <= 1e-4
            checked += 1
    assert checked >= 5


def test_serve_express_slot():
    # two slots: one takes requests in arrival order, the other (express) only
    # small ones. Two large grid requests queued ahead of a small one: the
    # first large one takes the ordinary slot, the small one the express slot
    # at once, and the second large one waits for the first; so the small
    # request is done before the second large one is admitted
    big1 = make_request(71, 30000, A.MODE_FFGRID)
    big2 = make_request(72, 30000, A.MODE_FFGRID)
    small = make_request(73, 1000, A.MODE_FFEXPLICIT)
    srv = A.Server(slots=2, budget_gb=100.0)
    for r in (big1, big2, small):  # resident first: the timed round has no host set-up
        srv.wait(srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))
    e0 = srv.stats()["express"]
    ids = [srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]) for r in (big1, big2, small)]
    (I1, l1, q1), (I2, l2, q2), (Is, ls, qs) = [srv.wait(i) for i in ids]
    st = srv.stats()
    print(f"big1 {l1:.2f} ms (queue {q1:.2f}), big2 {l2:.2f} ms (queue {q2:.2f}), small {ls:.2f} ms (queue {qs:.2f}),"
          f" express admissions {st['express'] - e0}")
    assert st["express"] - e0 == 1
    assert ls < q2, (ls, q2)
    assert np.array_equal(Is, direct(small))
    assert np.array_equal(I2, direct(big2))


def test_serve_errors():
    srv = A.Server(slots=2, budget_gb=1.0)
    r = make_request(5, 100, A.MODE_FFEXPLICIT)
    with pytest.raises(A.AusaxsError):
        srv.submit(r["mode"], r["xyz"], np.zeros(len(r["xyz"]), np.int32), r["wxyz"], r["q"])
    with pytest.raises(A.AusaxsError):
        srv.submit(A.MODE_SIMPLE, r["xyz"], r["ff"], r["wxyz"], r["q"])
    with pytest.raises(A.AusaxsError):
        A.Server(slots=0, budget_gb=1.0)


def test_serve_merged_launches():
    # small requests queued together run as merged launches (one engine, one
    # pair launch); each result equals the request alone, bit for bit
    reqs = [make_request(300 + k, n, MODES[k % 3]) for k, n in enumerate([250, 420, 333, 610, 180, 505, 777, 290])]
    srv = A.Server(slots=1, budget_gb=8.0)
    # a large request first keeps the one slot busy while the small ones queue
    big = make_request(299, 20000, A.MODE_FFEXPLICIT)
    ids = [srv.submit(big["mode"], big["xyz"], big["ff"], big["wxyz"], big["q"])]
    ids += [srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]) for r in reqs]
    res = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    print("stats:", st)
    assert st["completed"] == len(ids) and st["failed"] == 0
    assert st["batches"] >= 1 and st["batched"] >= 2
    for r, I in zip([big] + reqs, res):
        assert np.array_equal(I, direct(r)), r["mode"]


def test_serve_resident_structures():
    # a repeated structure is served from its resident copy: no rebuild and,
    # after the first request, no new device allocation at all
    r = make_request(41, 2500, A.MODE_FFGRID)
    srv = A.Server(slots=1, budget_gb=8.0)
    I0 = srv.wait(srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))[0]
    m0 = A.pool_stats()["device_mallocs"]
    for _ in range(6):
        I = srv.wait(srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))[0]
        assert np.array_equal(I, I0)
    st = srv.stats()
    print("stats:", st, "pool:", A.pool_stats())
    assert st["cache_hits"] == 6 and st["cache_misses"] == 1
    assert A.pool_stats()["device_mallocs"] == m0
    assert np.array_equal(I0, direct(r))
    # another mode of the same coordinates is another structure
    I_avg = srv.wait(srv.submit(A.MODE_FFAVG, r["xyz"], r["ff"], r["wxyz"], r["q"]))[0]
    assert srv.stats()["cache_misses"] == 2
    assert np.array_equal(I_avg, direct(dict(r, mode=A.MODE_FFAVG)))


def test_serve_steady_state_allocations():
    # a second pass over the same request mix allocates far less than the
    # first: resident structures and pooled buffers instead of hipMalloc
    reqs = requests(12, seed=9)
    srv = A.Server(slots=3, budget_gb=16.0)
    m0 = A.pool_stats()["device_mallocs"]
    first = [srv.wait(srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))[0] for r in reqs]
    m1 = A.pool_stats()["device_mallocs"]
    ids = [srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]) for r in reqs]
    second = [srv.wait(i)[0] for i in ids]
    m2 = A.pool_stats()["device_mallocs"]
    print("device mallocs: first pass", m1 - m0, "second pass", m2 - m1, "stats", srv.stats())
    for a, b in zip(first, second):
        assert np.array_equal(a, b)
    assert srv.stats()["cache_hits"] >= len(reqs) // 2
    assert m2 - m1 <= max(4, (m1 - m0) // 4)


def test_serve_eviction_under_budget():
    # resident structures give way to new requests when the budget is tight
    reqs = requests(10, seed=12)
    sizes = []
    for r in reqs:
        s1 = A.Server(slots=1, budget_gb=64.0)
        s1.wait(s1.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))
        sizes.append(s1.stats()["peak_gb"])
        s1.close()
    srv = A.Server(slots=2, budget_gb=1.2 * max(sizes))
    for rep in range(2):
        for r in reqs:
            I = srv.wait(srv.submit(r["mode"], r["xyz"], r["ff"], r["wxyz"], r["q"]))[0]
            assert np.array_equal(I, direct(r))
    st = srv.stats()
    print("stats:", st)
    assert st["failed"] == 0 and st["peak_gb"] <= 1.2 * max(sizes) + 1e-9
