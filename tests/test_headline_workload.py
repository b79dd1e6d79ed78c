"""The benchmark's headline workload at full size through the C ABI.

250,000-atom ellipsoid (semi-axes 120 x 90 x 70 A) with 60,000 waters,
FF-explicit histograms (HistogramManagerMTFFExplicit::calculate_all,
HistogramManagerMTFFExplicit.cpp:216-436) with 0.1 A weighted bins, I(q) on
1,000 q points up to 1.5 1/A with the bf16-split MFMA Debye transform. The
oracle runs the same input on the host's cores (about 6 s on 16 threads).

Bar: bin indices and counts bit-exact in every form-factor row, bin centres
within the bound of axis_ok (dense bins to a fraction of an ulp), and I(q)
within 1e-4 relative of the FP64 oracle at every q.
"""
import os

import numpy as np
import pytest

import ausaxs_amd as A
import oracle as O
from conftest import axis_ok

pytestmark = pytest.mark.gpu


def host_threads():
    """Every core this process may use, capped by OMP_NUM_THREADS where set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, min(n, 32))


@pytest.fixture(scope="module")
def workload():
    atoms = A.synth_globule(3, 250000, (120, 90, 70), 1.8)
    waters = A.synth_shell(4, 60000, (120, 90, 70), 3.0)
    q = np.linspace(0.0015, 1.5, 1000)
    ref = O.hist_ff(O.FFEXPLICIT, atoms, waters, 0.1, 8000, 1.0, host_threads())
    I_ref, _ = O.debye_ff(O.FFEXPLICIT, ref, q, 0.1)
    return atoms, waters, q, ref, I_ref


def test_250k_atoms_60k_waters_ffexplicit(workload):
    atoms, waters, q, ref, I_ref = workload
    st = A.Structure(A.MODE_FFEXPLICIT, atoms, waters, A.default_settings(bin_width=0.1))
    h = st.calculate()
    assert h.max_bin == ref["max_bin"]
    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww")):
        v, c, _ = h.get(which)
        assert np.array_equal(c, ref[key]["count"]), key
        np.testing.assert_allclose(v, ref[key]["value"], rtol=1e-15, err_msg=key)
    n, m = len(atoms), len(waters)
    assert h.get(A.AA)[1].sum(dtype=np.int64) == n * (n - 1) + n
    assert h.get(A.AW)[1].sum(dtype=np.int64) == n * m
    pt = ref["p_tot"]
    assert axis_ok(h.d_axis(), O.weighted_axis(pt, 0.1), 0.1, pt["count"])
    I = h.debye_transform(q)
    err = np.abs(I - I_ref) / np.abs(I_ref)
    print(f"I(q) max relative error {err.max():.3e} at q = {q[np.argmax(err)]:.4f}")
    assert err.max() <= 1e-4
    # a second calculate on the resident structure is bit-identical
    h2 = st.calculate()
    assert np.array_equal(h2.get(A.AA)[1], h.get(A.AA)[1])
    assert np.array_equal(h2.d_axis(), h.d_axis())
