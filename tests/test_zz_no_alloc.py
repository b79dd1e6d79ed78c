"""The allocation-free Debye contraction (ausaxs_debye_contract_noalloc)
against ausaxs_debye_contract, through the C ABI.

The caller owns the device buffers (allocated here with the HIP runtime, as a
caller under graph capture would have allocated them beforehand); the entry
point itself allocates nothing and works in a static device array. Two calls
in a row, the second with a smaller input, must give what the allocating entry
point gives (the same kernels and launch shapes: bit-identical), and a
contraction larger than the static work space must fail with an error before
any device work.
"""
import ctypes as C

import numpy as np
import pytest

import ausaxs_amd as A

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


@pytest.fixture(scope="module")
def hip():
    rt = C.CDLL("libamdhip64.so")
    rt.hipMalloc.restype = C.c_int
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.restype = C.c_int
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.restype = C.c_int
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.restype = C.c_int
    return rt


class DeviceArray:
    def __init__(self, rt, host):
        self.rt, self.n, self.p = rt, host.nbytes, C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(self.n, 8)) == 0
        if self.n:
            assert rt.hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def get(self, shape):
        out = np.zeros(shape, np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        assert self.rt.hipFree(self.p) == 0


def case(seed, rows, K, nq):
    rng = np.random.default_rng(seed)
    H = rng.integers(0, 10**6, (rows, K)).astype(np.float64)
    d = np.sort(rng.uniform(0.0, 250.0, K))
    d[0] = 0.0
    q = np.linspace(0.0015, 1.5, nq)
    return H, d, q


def noalloc(hip, H, d, q, terms):
    rows, K = H.shape
    nq = len(q)
    bufs = [DeviceArray(hip, np.ascontiguousarray(a)) for a in (H, d, q, np.zeros((rows, nq)))]
    try:
        A._check(A.lib.ausaxs_debye_contract_noalloc(bufs[0].p, rows, K, bufs[1].p, bufs[2].p, nq, terms, bufs[3].p,
                                                      None))
        assert hip.hipDeviceSynchronize() == 0
        return bufs[3].get((rows, nq))
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("terms", [3, 2])
def test_noalloc_matches_allocating_entry_point(hip, terms):
    # a 250k-atom-sized contraction (24 rows x 2461 bins x 1000 q), then a smaller one
    for seed, rows, K, nq in ((1, 24, 2461, 1000), (2, 5, 300, 77)):
        H, d, q = case(seed, rows, K, nq)
        ref = A.debye_contract(H, d, q, terms)
        out = noalloc(hip, H, d, q, terms)
        assert np.array_equal(out, ref), (seed, np.max(np.abs(out - ref)))


def test_noalloc_fp64_and_limits(hip):
    H, d, q = case(3, 7, 500, 120)
    assert np.array_equal(noalloc(hip, H, d, q, 0), A.debye_contract(H, d, q, 0))
    # 6 (nq + rows) K bytes of split operands alone exceed the static array
    H, d, q = case(4, 16, 8192, 4096)
    with pytest.raises(A.AusaxsError, match="work space"):
        noalloc(hip, H, d, q, 3)
