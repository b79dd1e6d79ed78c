"""The device's own issue ceilings (ausaxs_issue_probe) that the pair
kernel's roofline figures are quoted against."""
import math

import pytest

import ausaxs_amd as A


@pytest.mark.gpu
def test_issue_probe_rates():
    r = A.issue_probe(4926, 3)
    print("issue rates (wave instructions per CU per ns):", r)
    assert all(math.isfinite(v) and v > 0 for v in r.values())
    # gfx950 issues FP32 adds with only vector-register operands at about
    # twice the rate of the same add with a scalar-register operand, and the
    # transcendental at a quarter of it: the pair kernel keeps its constants
    # in vector registers because of the first
    assert r["fp32_add_vgpr"] > 1.3 * r["fp32_add_sgpr"]
    assert r["rsq"] < r["fp32_add_vgpr"]


@pytest.mark.gpu
def test_issue_probe_rejects_bad_arguments():
    with pytest.raises(A.AusaxsError):
        A.issue_probe(0, 1)
    with pytest.raises(A.AusaxsError):
        A.issue_probe(8193, 1)
    with pytest.raises(A.AusaxsError):
        A.issue_probe(4926, 0)
