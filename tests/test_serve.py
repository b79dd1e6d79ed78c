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


def test_serve_errors():
    srv = A.Server(slots=2, budget_gb=1.0)
    r = make_request(5, 100, A.MODE_FFEXPLICIT)
    with pytest.raises(A.AusaxsError):
        srv.submit(r["mode"], r["xyz"], np.zeros(len(r["xyz"]), np.int32), r["wxyz"], r["q"])
    with pytest.raises(A.AusaxsError):
        srv.submit(A.MODE_SIMPLE, r["xyz"], r["ff"], r["wxyz"], r["q"])
    with pytest.raises(A.AusaxsError):
        A.Server(slots=0, budget_gb=1.0)
