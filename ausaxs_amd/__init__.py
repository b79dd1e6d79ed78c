"""MI355X-native AUSAXS Debye hot path — Python binding of the C ABI.

The product is the C-ABI library ``libausaxs_hip.so`` (HIP kernels for gfx950
plus a C++20 host layer mirroring AUSAXS's histogram-manager API). This module
is a thin ctypes binding used by the tests and the benchmark. There is no CPU
fallback: loading fails loudly when the library is missing, and every call
fails loudly when no HIP device is present.

Reference interfaces mirrored (AUSAXS 1.2.9):
  * ``Structure.calculate`` -> ``IHistogramManager::calculate_all``
    (hist/histogram_manager/IHistogramManager.h:16-29)
  * ``CompositeHistogram.debye_transform`` -> ``ICompositeDistanceHistogram(Exv)::debye_transform``
    (hist/intensity_calculator/ICompositeDistanceHistogram.h:24-71)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libausaxs_hip.so"

MODE_FFAVG = 0
MODE_FFEXPLICIT = 1
MODE_SIMPLE = 2
MODE_FFGRID_SURFACE = 4
XX_INTERIOR, XX_SURFACE, XX_CROSS, WX_INTERIOR, WX_SURFACE, AX_INTERIOR, AX_SURFACE = 6, 7, 8, 9, 10, 11, 12
AA, AW, WW, PTOT = 0, 1, 2, 3
MAX_FF = 15

# form-factor slots (form_factor/FormFactorType.h:14-33)
FF = dict(EXV=0, WATER=1, OH=1, H=2, C=3, CH=4, CH2=5, CH3=6, N=7, NH=8, NH2=9, NH3=10, O=11, S=12, SH=13, OTHER=14)


class AusaxsError(RuntimeError):
    pass


class _Settings(C.Structure):
    _fields_ = [
        ("bin_width", C.c_double),
        ("bin_count", C.c_int),
        ("weighted_bins", C.c_int),
        ("debye_terms", C.c_int),
        ("tile_j", C.c_int),
        ("lattice", C.c_int),
        ("surface_thickness", C.c_double),
    ]


class _HydrateSettings(C.Structure):
    _fields_ = [
        ("atomic_radius", C.c_double),
        ("hydration_radius", C.c_double),
        ("shell_correction", C.c_double),
        ("explicit_grid", C.c_int),
        ("grid_min", C.c_double * 3),
        ("grid_max", C.c_double * 3),
    ]


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python ausaxs_amd/build.py` (hipcc, gfx950). "
            "There is no CPU fallback for the GPU path."
        )
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ausaxs_default_settings": (None, [P(_Settings)]),
        "ausaxs_get_last_error": (C.c_int, [C.c_char_p, C.c_int]),
        "ausaxs_device_info": (C.c_int, [C.c_char_p, C.c_int, P(C.c_int), P(C.c_size_t)]),
        "ausaxs_structure_create": (
            C.c_int,
            [C.c_int, vp, C.c_int, vp, C.c_int, P(_Settings), C.c_double, P(vp)],
        ),
        "ausaxs_structure_calculate": (C.c_int, [vp, P(vp)]),
        "ausaxs_structure_create_grid_surface": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, P(_Settings), P(vp)]),
        "ausaxs_structure_create_ffavg": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, P(_Settings), P(C.c_double),
                                                    P(vp)]),
        "ausaxs_hydrate_radial": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, P(C.c_int), vp,
                                            P(_HydrateSettings)]),
        "ausaxs_default_hydrate_settings": (None, [P(_HydrateSettings)]),
        "ausaxs_grid_exv_surface": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, P(C.c_int), vp, C.c_int,
                                              P(C.c_int), C.c_double]),
        "ausaxs_structure_free": (C.c_int, [vp]),
        "ausaxs_hist_calculate": (C.c_int, [C.c_int, vp, C.c_int, vp, C.c_int, P(_Settings), C.c_double, P(vp)]),
        "ausaxs_histogram_info": (C.c_int, [vp, P(C.c_int), P(C.c_int)]),
        "ausaxs_histogram_get": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "ausaxs_histogram_d_axis": (C.c_int, [vp, vp]),
        "ausaxs_histogram_set_params": (C.c_int, [vp] + [C.c_double] * 5),
        "ausaxs_histogram_debye": (C.c_int, [vp, vp, C.c_int, vp, vp]),
        "ausaxs_histogram_free": (C.c_int, [vp]),
        "ausaxs_default_q_axis": (C.c_int, [C.c_double, C.c_double, vp, P(C.c_int)]),
        "ausaxs_pair_histogram": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_double, C.c_int, vp, vp]),
        "ausaxs_debye_contract": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp]),
        "ausaxs_debye_contract_noalloc": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp]),
        "ausaxs_bin_check": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, vp]),
        "ausaxs_structure_set_timing": (C.c_int, [vp, C.c_int]),
        "ausaxs_copy_probe": (C.c_int, [C.c_size_t, C.c_int, P(C.c_double)]),
        "ausaxs_issue_probe": (C.c_int, [C.c_int, C.c_int, P(C.c_double)]),
        "ausaxs_rigid_create": (C.c_int, [vp, vp, vp, C.c_int, vp, vp, C.c_int, vp, P(vp)]),
        "ausaxs_rigid_scan": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, C.c_double, vp]),
        "ausaxs_rigid_scan_hydrated": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_double, vp]),
        "ausaxs_rigid_histogram": (C.c_int, [vp, C.c_int, P(vp)]),
        "ausaxs_rigid_info": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int),
                                        P(C.c_longlong), P(C.c_double)]),
        "ausaxs_rigid_free": (C.c_int, [vp]),
        "ausaxs_server_create": (C.c_int, [C.c_int, C.c_double, P(vp)]),
        "ausaxs_server_submit": (C.c_int, [vp, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_double,
                                           P(C.c_longlong)]),
        "ausaxs_server_wait": (C.c_int, [vp, C.c_longlong, C.c_double, vp, P(C.c_double), P(C.c_double),
                                         P(C.c_int)]),
        "ausaxs_server_stats": (C.c_int, [vp, P(C.c_longlong), P(C.c_longlong), P(C.c_longlong), P(C.c_int),
                                          P(C.c_int), P(C.c_double), P(C.c_longlong), P(C.c_longlong)]),
        "ausaxs_server_free": (C.c_int, [vp]),
        "ausaxs_server_batch_stats": (C.c_int, [vp] + [P(C.c_longlong)] * 5 + [P(C.c_double), P(C.c_longlong)]),
        "ausaxs_pool_stats": (C.c_int, [P(C.c_size_t), P(C.c_size_t), P(C.c_longlong), P(C.c_longlong),
                                        P(C.c_longlong)]),
        "ausaxs_pool_trim": (C.c_int, []),
        "ausaxs_lsq": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp]),
        "ausaxs_cubic_spline": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp]),
        "ausaxs_exact_debye": (C.c_int, [vp, C.c_int, vp, C.c_int, vp]),
        "ausaxs_last_timing": (C.c_int, [vp, P(C.c_double), P(C.c_double), P(C.c_longlong), P(C.c_int),
                                         P(C.c_longlong), P(C.c_longlong)]),
        "ausaxs_last_merge_timing": (C.c_int, [vp, P(C.c_double), P(C.c_longlong)]),
        "ausaxs_fitter_create": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_uint, vp, C.c_int, P(vp)]),
        "ausaxs_scalable_create": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, P(_Settings), P(vp)]),
        "ausaxs_scalable_calculate": (C.c_int, [vp, C.c_double, P(vp)]),
        "ausaxs_scalable_free": (C.c_int, [vp]),
        "ausaxs_fitter_create_scalable": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_uint, vp, C.c_int, P(vp)]),
        "ausaxs_fitter_params": (C.c_int, [vp, P(C.c_int), P(C.c_uint)]),
        "ausaxs_fitter_evaluate": (C.c_int, [vp, vp, C.c_int, vp]),
        "ausaxs_fitter_fit": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, P(C.c_int), P(C.c_int), vp, vp]),
        "ausaxs_fitter_free": (C.c_int, [vp]),
        "ausaxs_symmetry_create": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, C.c_int, P(_Settings),
                                             P(vp)]),
        "ausaxs_symmetry_create_partial": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, C.c_int,
                                                     P(_Settings), P(vp)]),
        "ausaxs_symmetry_set_body": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "ausaxs_symmetry_set_symmetry": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp]),
        "ausaxs_symmetry_set_waters": (C.c_int, [vp, vp, C.c_int]),
        "ausaxs_symmetry_calculate": (C.c_int, [vp, P(vp)]),
        "ausaxs_symmetry_free": (C.c_int, [vp]),
        "ausaxs_symmetry_info": (C.c_int, [vp, P(C.c_longlong), P(C.c_int)]),
        "ausaxs_symmetry_chain_schedule": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, P(C.c_int)]),
        "ausaxs_symmetry_placement_schedule": (C.c_int, [vp, C.c_int, vp, C.c_int, P(C.c_int)]),
        "ausaxs_symmetry_cyclic": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, vp, P(C.c_int)]),
        "ausaxs_synth_globule": (
            C.c_int,
            [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, vp],
        ),
        "ausaxs_synth_shell": (
            C.c_int,
            [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, vp],
        ),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int):
    if rc != 0:
        buf = C.create_string_buffer(4096)
        lib.ausaxs_get_last_error(buf, 4096)
        raise AusaxsError(buf.value.decode(errors="replace"))


def default_settings(**kw) -> _Settings:
    s = _Settings()
    lib.ausaxs_default_settings(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def device_info():
    name = C.create_string_buffer(256)
    cus, lds = C.c_int(), C.c_size_t()
    _check(lib.ausaxs_device_info(name, 256, C.byref(cus), C.byref(lds)))
    return dict(arch=name.value.decode(), compute_units=cus.value, lds_per_block=lds.value)


def points_ff(xyz: np.ndarray, ff) -> np.ndarray:
    """Pack FP32 coordinates and form-factor types into the XYZFF layout."""
    xyz = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((len(xyz), 4), dtype=np.float32)
    out[:, :3] = xyz
    out.view(np.int32)[:, 3] = np.broadcast_to(np.asarray(ff, dtype=np.int32), (len(xyz),))
    return out


def points_w(xyz: np.ndarray, w) -> np.ndarray:
    """Pack FP32 coordinates and weights into the XYZW layout."""
    xyz = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((len(xyz), 4), dtype=np.float32)
    out[:, :3] = xyz
    out[:, 3] = np.broadcast_to(np.asarray(w, dtype=np.float32), (len(xyz),))
    return out


def _pts(a):
    if a is None:
        return np.zeros((0, 4), np.float32)
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 4:
        raise ValueError("points must be an (n, 4) float32 array")
    return a


class CompositeHistogram:
    """ICompositeDistanceHistogram(Exv) returned by calculate_all."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        mode, mb = C.c_int(), C.c_int()
        _check(lib.ausaxs_histogram_info(self._h, C.byref(mode), C.byref(mb)))
        self.mode, self.max_bin = mode.value, mb.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.ausaxs_histogram_free(h)
            self._h = C.c_void_p(0)

    def _shape(self, which):
        mb = self.max_bin
        if self.mode == MODE_SIMPLE or which in (WW, PTOT, 4, 5, 6, 7, 8, 9, 10):  # FFGrid / surface 1-D profiles
            return (mb,)
        return (MAX_FF, MAX_FF, mb) if which == AA else (MAX_FF, mb)

    def get(self, which):
        """(value, count, bin_center) arrays in the reference layout."""
        shp = self._shape(which)
        v = np.zeros(shp, np.float64)
        c = np.zeros(shp, np.uint32)
        b = np.zeros(shp, np.float64)
        _check(lib.ausaxs_histogram_get(self._h, which, _ptr(v), _ptr(c), _ptr(b)))
        return v, c, b

    def d_axis(self):
        d = np.zeros(self.max_bin, np.float64)
        _check(lib.ausaxs_histogram_d_axis(self._h, _ptr(d)))
        return d

    def set_params(self, cw=1.0, cx=1.0, crho=1.0, dw_atomic=0.0, dw_exv=0.0):
        _check(lib.ausaxs_histogram_set_params(self._h, cw, cx, crho, dw_atomic, dw_exv))

    def debye_transform(self, q, profiles=False):
        q = np.ascontiguousarray(q, dtype=np.float64)
        I = np.zeros(len(q), np.float64)
        prof = np.zeros((6, len(q)), np.float64) if profiles else None
        _check(lib.ausaxs_histogram_debye(self._h, _ptr(q), len(q), _ptr(I), _ptr(prof)))
        return (I, prof) if profiles else I


class Structure:
    """Resident coordinates + work lists; calculate() = calculate_all()."""

    def __init__(self, mode, atoms, waters=None, settings=None, z_exv=1.0):
        self._atoms = _pts(atoms)
        self._waters = _pts(waters)
        s = settings or default_settings()
        h = C.c_void_p()
        _check(
            lib.ausaxs_structure_create(
                mode, _ptr(self._atoms), len(self._atoms), _ptr(self._waters), len(self._waters),
                C.byref(s), float(z_exv), C.byref(h),
            )
        )
        self._s = h

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib.ausaxs_structure_free(s)
            self._s = C.c_void_p(0)

    def set_timing(self, on=True):
        _check(lib.ausaxs_structure_set_timing(self._s, int(on)))

    def calculate(self) -> CompositeHistogram:
        h = C.c_void_p()
        _check(lib.ausaxs_structure_calculate(self._s, C.byref(h)))
        return CompositeHistogram(h.value)

    def last_timing(self):
        pm, tm, pr, ni = C.c_double(), C.c_double(), C.c_longlong(), C.c_int()
        kp, lp = C.c_longlong(), C.c_longlong()
        _check(lib.ausaxs_last_timing(self._s, C.byref(pm), C.byref(tm), C.byref(pr), C.byref(ni), C.byref(kp),
                                      C.byref(lp)))
        mm, mb = C.c_double(), C.c_longlong()
        _check(lib.ausaxs_last_merge_timing(self._s, C.byref(mm), C.byref(mb)))
        return dict(pair_ms=pm.value, total_ms=tm.value, pairs=pr.value, n_items=ni.value, kernel_pairs=kp.value,
                    lattice_pairs=lp.value, merge_ms=mm.value, merge_bytes=mb.value)


def calculate_all(mode, atoms, waters=None, settings=None, z_exv=1.0) -> CompositeHistogram:
    return Structure(mode, atoms, waters, settings, z_exv).calculate()


def default_q_axis(qmin=1e-4, qmax=0.5):
    q = np.zeros(200, np.float64)
    n = C.c_int()
    _check(lib.ausaxs_default_q_axis(qmin, qmax, _ptr(q), C.byref(n)))
    return q[: n.value].copy()


def pair_histogram(a, b=None, bin_width=0.25, bin_count=8000):
    """Raw pair histogram of one set (reference loop geometry): (count, sum of distances)."""
    a = _pts(a)
    bb = None if b is None else _pts(b)
    cnt = np.zeros(bin_count, np.uint64)
    ds = np.zeros(bin_count, np.float64)
    _check(
        lib.ausaxs_pair_histogram(
            _ptr(a), len(a), _ptr(bb), 0 if bb is None else len(bb), bin_width, bin_count, _ptr(cnt), _ptr(ds)
        )
    )
    return cnt, ds


def debye_contract(H, d, q, terms=3):
    H = np.ascontiguousarray(H, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    rows, K = H.shape
    out = np.zeros((rows, len(q)), np.float64)
    _check(lib.ausaxs_debye_contract(_ptr(H), rows, K, _ptr(d), _ptr(q), len(q), terms, _ptr(out)))
    return out


def bin_check(lo_bits, hi_bits, bin_width):
    """Exhaustive device check over float d2 bit patterns [lo_bits, hi_bits]:
    (bins differing from the correctly rounded rule, device sqrt results not
    correctly rounded, tie-fallback corrections differing from it, offset codes
    more than one unit from the real root's, mean code error against the real
    root in units, the same mean for the correctly rounded root)."""
    m = np.zeros(6, np.uint64)
    _check(lib.ausaxs_bin_check(lo_bits, hi_bits, bin_width, _ptr(m)))
    n = hi_bits - lo_bits + 1
    mean = [float(np.int64(np.uint64(v).view(np.int64))) / 65536.0 / n for v in (m[4], m[5])]
    return int(m[0]), int(m[1]), int(m[2]), int(m[3]), mean[0], mean[1]


def issue_probe(nbins=4926, reps=5):
    """The device's issue ceilings for the pair kernel's instruction classes,
    in wave instructions per CU per ns: FP32 adds with vector-register
    operands, with a scalar-register operand, v_rsq_f32, and ds_add_u64 at
    hashed cells of an nbins-cell LDS histogram."""
    r = (C.c_double * 4)()
    _check(lib.ausaxs_issue_probe(int(nbins), int(reps), r))
    return {"fp32_add_vgpr": r[0], "fp32_add_sgpr": r[1], "rsq": r[2], "lds_add_u64": r[3]}


def synth_globule(seed, n, axes, r_excl):
    out = np.zeros((n, 4), np.float32)
    _check(lib.ausaxs_synth_globule(seed, n, *map(float, axes), float(r_excl), _ptr(out)))
    return out


def synth_shell(seed, m, axes, offset=3.0):
    out = np.zeros((m, 4), np.float32)
    _check(lib.ausaxs_synth_shell(seed, m, *map(float, axes), float(offset), _ptr(out)))
    return out


# ---------------------------------------------------------------------------
# grid excluded volume (HistogramManagerMTFFGrid)
# ---------------------------------------------------------------------------
MODE_FFGRID = 3
PTOT_AX, XX = 4, 5
_P = C.POINTER
lib.ausaxs_structure_create_grid.restype = C.c_int
lib.ausaxs_structure_create_grid.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, _P(_Settings),
                                             _P(C.c_void_p)]
lib.ausaxs_grid_exv.restype = C.c_int
lib.ausaxs_grid_exv.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                _P(C.c_int), _P(C.c_double)]
lib.ausaxs_histogram_grid_axes.restype = C.c_int
lib.ausaxs_histogram_grid_axes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]


def grid_exv(atoms_xyz, ff, waters_xyz=None):
    """Excluded-volume dummies (reference order) and grid volume; host-only."""
    a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ff, np.int32).reshape(-1)
    w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)), np.float64).reshape(-1, 3)
    n, vol = C.c_int(), C.c_double()
    rc = lib.ausaxs_grid_exv(_ptr(a), _ptr(f), len(a), _ptr(w), len(w), None, 0, C.byref(n), C.byref(vol))
    out = np.zeros((max(1, n.value), 3), np.float64)
    _check(lib.ausaxs_grid_exv(_ptr(a), _ptr(f), len(a), _ptr(w), len(w), _ptr(out), n.value, C.byref(n),
                               C.byref(vol)))
    return out[: n.value].copy(), vol.value


class FFAvgStructure(Structure):
    """HistogramManagerMTFFAvg from FP64 atoms and waters, with Z_exv from the
    molecule's grid volume as the reference manager computes it."""

    def __init__(self, atoms_xyz, ff, waters_xyz=None, settings=None):
        a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
        f = np.ascontiguousarray(ff, np.int32).reshape(-1)
        w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)), np.float64).reshape(-1, 3)
        s = settings or default_settings()
        h, z = C.c_void_p(), C.c_double()
        _check(lib.ausaxs_structure_create_ffavg(_ptr(a), _ptr(f), len(a), _ptr(w), len(w), C.byref(s), C.byref(z),
                                                 C.byref(h)))
        self._s = h
        self.z_exv = z.value


class GridStructure(Structure):
    """HistogramManagerMTFFGrid on resident data (FP64 input coordinates)."""

    def __init__(self, atoms_xyz, ff, waters_xyz=None, settings=None):
        self._a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
        self._f = np.ascontiguousarray(ff, np.int32).reshape(-1)
        self._w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)),
                                       np.float64).reshape(-1, 3)
        s = settings or default_settings()
        h = C.c_void_p()
        _check(lib.ausaxs_structure_create_grid(_ptr(self._a), _ptr(self._f), len(self._a), _ptr(self._w),
                                                len(self._w), C.byref(s), C.byref(h)))
        self._s = h


class GridSurfaceStructure(Structure):
    """HistogramManagerMTFFGridSurface on resident data: the grid dummies split
    into interior and surface ones; set_params(cx=...) scales the surface."""

    def __init__(self, atoms_xyz, ff, waters_xyz=None, settings=None):
        self._a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
        self._f = np.ascontiguousarray(ff, np.int32).reshape(-1)
        self._w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)),
                                       np.float64).reshape(-1, 3)
        s = settings or default_settings()
        h = C.c_void_p()
        _check(lib.ausaxs_structure_create_grid_surface(_ptr(self._a), _ptr(self._f), len(self._a), _ptr(self._w),
                                                        len(self._w), C.byref(s), C.byref(h)))
        self._s = h


def hydrate(atoms_xyz, ff, bodies=None, noise=None, atomic_radius=None, hydration_radius=None,
            shell_correction=None, grid_limits=None):
    """Radial hydration shell (RadialHydration with NoCulling); host-only.

    bodies: body sizes (None: one body). noise: None (no displacement, as the
    reference's own tests use), an int seed (N(0, 0.75) per component, the
    reference's default distribution, from numpy), or an (n, 26, 3) array.
    atomic_radius / hydration_radius / shell_correction: the grid's radii and
    settings::hydrate::shell_correction (None: the reference's defaults);
    grid_limits: ((xmin, ymin, zmin), (xmax, ymax, zmax)) for a grid of
    Grid(Limit3D) instead of the one built from the atoms.
    Returns (waters n x 3 FP64 in placement order, waters per body)."""
    hs = _HydrateSettings()
    lib.ausaxs_default_hydrate_settings(C.byref(hs))
    if atomic_radius is not None:
        hs.atomic_radius = float(atomic_radius)
    if hydration_radius is not None:
        hs.hydration_radius = float(hydration_radius)
    if shell_correction is not None:
        hs.shell_correction = float(shell_correction)
    if grid_limits is not None:
        hs.explicit_grid = 1
        for k in range(3):
            hs.grid_min[k] = float(grid_limits[0][k])
            hs.grid_max[k] = float(grid_limits[1][k])
    a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ff, np.int32).reshape(-1)
    n = len(a)
    b = np.ascontiguousarray([n] if bodies is None else bodies, np.int32)
    if isinstance(noise, (int, np.integer)) and not isinstance(noise, bool):
        noise = np.random.default_rng(int(noise)).normal(0.0, 0.75, (n, 26, 3))
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64).reshape(n, 26, 3)
    out = np.zeros((max(1, 26 * n), 3), np.float64)
    m = C.c_int()
    per = np.zeros(len(b), np.int32)
    _check(lib.ausaxs_hydrate_radial(_ptr(a), _ptr(f), n, _ptr(b), len(b), None if nz is None else _ptr(nz),
                                     _ptr(out), len(out), C.byref(m), _ptr(per), C.byref(hs)))
    return out[: m.value].copy(), per


def grid_exv_surface(atoms_xyz, ff, waters_xyz=None, surface_thickness=1.0):
    """Interior and surface dummies (RawGridWithSurfaceExv, reference order); host-only.
    surface_thickness: settings::grid::exv::surface_thickness in A (cell width 1)."""
    a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ff, np.int32).reshape(-1)
    w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)), np.float64).reshape(-1, 3)
    ni, ns = C.c_int(), C.c_int()
    t = C.c_double(surface_thickness)
    _check(lib.ausaxs_grid_exv_surface(_ptr(a), _ptr(f), len(a), _ptr(w), len(w), None, 0, C.byref(ni), None, 0,
                                       C.byref(ns), t))
    xi, xs = np.zeros((max(1, ni.value), 3)), np.zeros((max(1, ns.value), 3))
    _check(lib.ausaxs_grid_exv_surface(_ptr(a), _ptr(f), len(a), _ptr(w), len(w), _ptr(xi), ni.value, C.byref(ni),
                                       _ptr(xs), ns.value, C.byref(ns), t))
    return xi[: ni.value].copy(), xs[: ns.value].copy()


def _grid_axes(self):
    d_ax = np.zeros(self.max_bin)
    d_xx = np.zeros(self.max_bin)
    _check(lib.ausaxs_histogram_grid_axes(self._h, _ptr(d_ax), _ptr(d_xx)))
    return d_ax, d_xx


CompositeHistogram.grid_axes = _grid_axes


class RigidScan:
    """PartialHistogramManagerMT over many rigid-body conformations (one hipGraph per batch).

    bodies: list of (n_b x 3) FP64 coordinate arrays, weights: list of per-atom
    weights (XYZW). Self terms are cached from these coordinates; scan() moves
    every body by its transform T[c, b] = (R row-major, t) per conformation."""

    def __init__(self, bodies, weights, waters_xyz=None, waters_w=None, settings=None, autotune=False):
        self.sizes = np.array([len(b) for b in bodies], np.int32)
        self._xyz = np.ascontiguousarray(np.concatenate([np.asarray(b, np.float64).reshape(-1, 3) for b in bodies])
                                         if len(bodies) else np.zeros((0, 3)), np.float64)
        self._w = np.ascontiguousarray(np.concatenate([np.broadcast_to(np.asarray(w, np.float32), (len(b),))
                                                       for b, w in zip(bodies, weights)]), np.float32)
        m = 0 if waters_xyz is None else len(waters_xyz)
        self._wx = np.ascontiguousarray(waters_xyz if m else np.zeros((0, 3)), np.float64).reshape(-1, 3)
        self._ww = np.ascontiguousarray(np.broadcast_to(np.asarray(waters_w if m else 1.0, np.float32), (m,)),
                                        np.float32)
        s = settings or default_settings()
        if autotune:
            s.tile_j = 0
        h = C.c_void_p()
        _check(lib.ausaxs_rigid_create(_ptr(self._xyz), _ptr(self._w), _ptr(self.sizes), len(self.sizes),
                                       _ptr(self._wx), _ptr(self._ww), m, C.byref(s), C.byref(h)))
        self._r = h

    def __del__(self):
        if getattr(self, "_r", None):
            lib.ausaxs_rigid_free(self._r)
            self._r = None

    def scan(self, T, q, cw=1.0):
        T = np.ascontiguousarray(T, np.float64).reshape(-1, len(self.sizes), 12)
        q = np.ascontiguousarray(q, np.float64)
        I = np.zeros((len(T), len(q)))
        _check(lib.ausaxs_rigid_scan(self._r, _ptr(T), len(T), _ptr(q), len(q), cw, _ptr(I)))
        return I

    def scan_hydrated(self, T, waters, q, cw=1.0):
        """Every conformation with its own waters: waters[c] = (xyz (m x 3), weights (m or scalar))."""
        T = np.ascontiguousarray(T, np.float64).reshape(-1, len(self.sizes), 12)
        if len(waters) != len(T):
            raise ValueError("one water set per conformation")
        counts = np.array([len(np.asarray(x).reshape(-1, 3)) for x, _ in waters], np.int32)
        wx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1, 3) for x, _ in waters])
                                  if len(waters) else np.zeros((0, 3)), np.float64)
        ww = np.ascontiguousarray(np.concatenate([np.broadcast_to(np.asarray(w, np.float32), (n,))
                                                  for (_, w), n in zip(waters, counts)])
                                  if len(waters) else np.zeros(0), np.float32)
        q = np.ascontiguousarray(q, np.float64)
        I = np.zeros((len(T), len(q)))
        _check(lib.ausaxs_rigid_scan_hydrated(self._r, _ptr(T), len(T), _ptr(wx), _ptr(ww), _ptr(counts), _ptr(q),
                                              len(q), cw, _ptr(I)))
        return I

    def histogram(self, c) -> CompositeHistogram:
        h = C.c_void_p()
        _check(lib.ausaxs_rigid_histogram(self._r, c, C.byref(h)))
        return CompositeHistogram(h.value)

    def info(self):
        v = [C.c_int() for _ in range(5)]
        pairs, ms = C.c_longlong(), C.c_double()
        _check(lib.ausaxs_rigid_info(self._r, *map(C.byref, v), C.byref(pairs), C.byref(ms)))
        return dict(tile_j=v[0].value, grid=v[1].value, copies=v[2].value, batch=v[3].value, n_items=v[4].value,
                    pairs=pairs.value, scan_ms=ms.value)


FIT_HYDRATION, FIT_EXCLUDED_VOLUME, FIT_SOLVENT_DENSITY, FIT_ATOMIC_DW, FIT_EXV_DW = 1, 2, 4, 8, 16
MINI_GOLDEN, MINI_SCAN = 0, 2


class ScalableGridStructure:
    """HistogramManagerMTFFGridScalableExv: the grid structure whose dummies
    can be stretched; calculate(k) is the FFGrid histogram at stretch k."""

    def __init__(self, atoms_xyz, ff, waters_xyz=None, settings=None):
        self._xyz = np.ascontiguousarray(np.asarray(atoms_xyz, np.float64).reshape(-1, 3))
        self._ff = np.ascontiguousarray(np.asarray(ff, np.int32).reshape(-1))
        self._w = np.ascontiguousarray(np.zeros((0, 3)) if waters_xyz is None else
                                       np.asarray(waters_xyz, np.float64).reshape(-1, 3))
        s = settings or default_settings()
        h = C.c_void_p()
        _check(lib.ausaxs_scalable_create(_ptr(self._xyz), _ptr(self._ff), len(self._xyz), _ptr(self._w),
                                          len(self._w), C.byref(s), C.byref(h)))
        self._s = h

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib.ausaxs_scalable_free(s)
            self._s = C.c_void_p(0)

    def calculate(self, k=1.0) -> CompositeHistogram:
        h = C.c_void_p()
        _check(lib.ausaxs_scalable_calculate(self._s, float(k), C.byref(h)))
        return CompositeHistogram(h.value)


class Fitter:
    """SmartFitter(data, histogram): fit of a composite histogram to measured
    I(q). The histogram (or scalable grid structure) must stay alive while the
    fitter is used."""

    def __init__(self, hist, q, I, err=None, enabled=FIT_HYDRATION, q_model=None):
        self._hist = hist
        self.q = np.ascontiguousarray(q, np.float64)
        self.I = np.ascontiguousarray(I, np.float64)
        self.err = None if err is None else np.ascontiguousarray(err, np.float64)
        qm = None if q_model is None else np.ascontiguousarray(q_model, np.float64)
        h = C.c_void_p()
        make = lib.ausaxs_fitter_create_scalable if isinstance(hist, ScalableGridStructure) else \
            lib.ausaxs_fitter_create
        handle = hist._s if isinstance(hist, ScalableGridStructure) else hist._h
        _check(make(handle, _ptr(self.q), _ptr(self.I), _ptr(self.err), len(self.q),
                    int(enabled), _ptr(qm), 0 if qm is None else len(qm), C.byref(h)))
        self._f = h
        k, en = C.c_int(), C.c_uint()
        _check(lib.ausaxs_fitter_params(self._f, C.byref(k), C.byref(en)))
        self.n_params, self.enabled = k.value, en.value

    def __del__(self):
        f = getattr(self, "_f", None)
        if f is not None and f.value:
            lib.ausaxs_fitter_free(f)
            self._f = C.c_void_p(0)

    def evaluate(self, params):
        """(n_sets, 5) array of a, b, var a, var b, chi2 for (n_sets, n_params) parameters."""
        p = np.ascontiguousarray(np.asarray(params, np.float64).reshape(-1, max(self.n_params, 1))
                                 if self.n_params else np.zeros((np.size(params) or 1, 0)), np.float64)
        n_sets = len(p)
        out = np.zeros((n_sets, 5), np.float64)
        _check(lib.ausaxs_fitter_evaluate(self._f, _ptr(p), n_sets, _ptr(out)))
        return out

    def fit(self, algorithm=MINI_GOLDEN, scan_evals=100):
        n = len(self.q)
        params, perr = np.zeros(max(self.n_params, 1)), np.zeros(max(self.n_params, 1))
        lin = np.zeros(5)
        dof, fevals = C.c_int(), C.c_int()
        I_fit, res = np.zeros(n), np.zeros(n)
        _check(lib.ausaxs_fitter_fit(self._f, int(algorithm), int(scan_evals), _ptr(params), _ptr(perr), _ptr(lin),
                                     C.byref(dof), C.byref(fevals), _ptr(I_fit), _ptr(res)))
        return dict(params=params[: self.n_params], param_err=perr[: self.n_params], a=lin[0], b=lin[1],
                    a_err=lin[2], b_err=lin[3], chi2=lin[4], dof=dof.value, fevals=fevals.value, I_fit=I_fit,
                    residuals=res)


class Symmetry:
    """One symmetry of a body: `transforms` (reps x 12 doubles: rotation
    row-major, translation) generate copies 1..reps; `schedule` lists the
    distinct copy pairs (repA, repB, scale), repetition 0 being the original
    (ISymmetry::internal_pair_schedule)."""

    def __init__(self, transforms, schedule):
        self.transforms = np.ascontiguousarray(np.asarray(transforms, np.float64).reshape(-1, 12))
        self.schedule = np.ascontiguousarray(np.asarray(schedule, np.int32).reshape(-1, 3))

    @property
    def reps(self):
        return len(self.transforms)

    @classmethod
    def cyclic(cls, cm, offset, step, axis, angle, reps):
        """CyclicSymmetry(offset, step, axis, angle, reps) of a body with
        centre cm, with its chain schedule."""
        out = np.zeros((reps, 12), np.float64)
        closed = C.c_int()
        v = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(3)) for x in (cm, offset, step, axis)]
        _check(lib.ausaxs_symmetry_cyclic(*[_ptr(x) for x in v], float(angle), int(reps), _ptr(out),
                                          C.byref(closed)))
        return cls(out, chain_schedule(reps, bool(closed.value)))


def _schedule_out(call):
    n = C.c_int()
    _check(call(None, 0, C.byref(n)))
    out = np.zeros((max(n.value, 1), 3), np.int32)
    _check(call(_ptr(out), n.value, C.byref(n)))
    return out[: n.value]


def chain_schedule(reps, closed):
    return _schedule_out(lambda o, c, n: lib.ausaxs_symmetry_chain_schedule(int(reps), int(bool(closed)), o, c, n))


def placement_schedule(placements):
    p = np.ascontiguousarray(np.asarray(placements, np.float64).reshape(-1, 12))
    return _schedule_out(lambda o, c, n: lib.ausaxs_symmetry_placement_schedule(_ptr(p), len(p), o, c, n))


class SymmetryStructure:
    """Bodies with symmetry copies; calculate() = SymmetryManagerMT::calculate_all.

    bodies: list of (n_b, 4) float arrays {x, y, z, w}; symmetries: list (one
    per body) of lists of Symmetry; waters: (m, 4) floats or None."""

    def __init__(self, bodies, symmetries, waters=None, settings=None, partial=False):
        """partial=True: PartialSymmetryManagerMT (its pair roles, self terms
        kept from the last internal change; see set_body)."""
        bodies = [_pts(b) for b in bodies]
        self._n = [len(b) for b in bodies]
        if len(symmetries) != len(bodies):
            raise ValueError("one symmetry list per body")
        self._atoms = np.ascontiguousarray(np.concatenate(bodies) if bodies else np.zeros((0, 4), np.float32))
        self._sizes = np.array([len(b) for b in bodies], np.int32)
        syms = [s for per in symmetries for s in per]
        self._per = np.array([len(per) for per in symmetries], np.int32)
        self._reps = np.array([s.reps for s in syms], np.int32)
        self._T = np.ascontiguousarray(np.concatenate([s.transforms for s in syms]) if syms
                                       else np.zeros((0, 12)), np.float64)
        self._slen = np.array([len(s.schedule) for s in syms], np.int32)
        self._sched = np.ascontiguousarray(np.concatenate([s.schedule for s in syms]) if syms
                                           else np.zeros((0, 3), np.int32), np.int32)
        self._waters = _pts(waters)
        s = settings or default_settings()
        h = C.c_void_p()
        create = lib.ausaxs_symmetry_create_partial if partial else lib.ausaxs_symmetry_create
        _check(create(
            _ptr(self._atoms), _ptr(self._sizes), len(self._sizes), _ptr(self._per), _ptr(self._reps),
            _ptr(self._T), _ptr(self._slen), _ptr(self._sched), _ptr(self._waters), len(self._waters),
            C.byref(s), C.byref(h)))
        self._s = h

    def set_body(self, b, points, internal=False):
        """New float XYZW coordinates of body b (same size): a rigid move, or
        with internal=True a change of its own distances or weights."""
        p = _pts(points)
        if len(p) != self._n[b]:
            raise ValueError("a body keeps its size")
        _check(lib.ausaxs_symmetry_set_body(self._s, int(b), _ptr(p), int(bool(internal))))

    def set_symmetry(self, b, k, sym):
        _check(lib.ausaxs_symmetry_set_symmetry(self._s, int(b), int(k), sym.reps, _ptr(sym.transforms),
                                                len(sym.schedule), _ptr(sym.schedule)))

    def set_waters(self, waters):
        w = _pts(waters)
        _check(lib.ausaxs_symmetry_set_waters(self._s, _ptr(w), len(w)))

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib.ausaxs_symmetry_free(s)
            self._s = C.c_void_p(0)

    def calculate(self) -> CompositeHistogram:
        h = C.c_void_p()
        _check(lib.ausaxs_symmetry_calculate(self._s, C.byref(h)))
        return CompositeHistogram(h.value)

    def info(self):
        pairs, items = C.c_longlong(), C.c_int()
        _check(lib.ausaxs_symmetry_info(self._s, C.byref(pairs), C.byref(items)))
        return dict(pairs=pairs.value, n_items=items.value)


def rigid_transform(R, t):
    """12-entry transform (R row-major, t)."""
    return np.concatenate([np.asarray(R, np.float64).reshape(9), np.asarray(t, np.float64).reshape(3)])


def random_rotation(rng):
    """Uniform random rotation (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(3)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    x, y, z, w = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3), \
        b * np.cos(2 * np.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


class Server:
    """Request loop with continuous batching over `slots` device slots and a
    device-memory budget (GiB). submit() returns an id; wait() returns
    (I, latency_ms, queue_ms) or None on timeout."""

    def __init__(self, slots=4, budget_gb=64.0):
        h = C.c_void_p()
        _check(lib.ausaxs_server_create(slots, budget_gb, C.byref(h)))
        self._s = h
        self._nq = {}

    def __del__(self):
        if getattr(self, "_s", None):
            lib.ausaxs_server_free(self._s)
            self._s = None

    def close(self):
        self.__del__()

    def submit(self, mode, atoms_xyz, ff, waters_xyz=None, q=None, settings=None, z_exv=1.0):
        a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(ff, np.int32), (len(a),)), np.int32)
        w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)), np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q if q is not None else default_q_axis(), np.float64)
        s = settings or default_settings()
        rid = C.c_longlong()
        _check(lib.ausaxs_server_submit(self._s, mode, _ptr(a), _ptr(f), len(a), _ptr(w), len(w), _ptr(q), len(q),
                                        C.byref(s), z_exv, C.byref(rid)))
        self._nq[rid.value] = len(q)
        return rid.value

    def wait(self, rid, timeout=600.0):
        I = np.zeros(self._nq[rid])
        lat, qu, ready = C.c_double(), C.c_double(), C.c_int()
        try:
            _check(lib.ausaxs_server_wait(self._s, rid, timeout, _ptr(I), C.byref(lat), C.byref(qu), C.byref(ready)))
        except AusaxsError:
            self._nq.pop(rid, None)
            raise
        if not ready.value:
            return None
        self._nq.pop(rid, None)
        return I, lat.value, qu.value

    def stats(self):
        v = [C.c_longlong() for _ in range(3)]
        fl, pk, gb, pairs, kp = C.c_int(), C.c_int(), C.c_double(), C.c_longlong(), C.c_longlong()
        _check(lib.ausaxs_server_stats(self._s, *map(C.byref, v), C.byref(fl), C.byref(pk), C.byref(gb),
                                       C.byref(pairs), C.byref(kp)))
        b = [C.c_longlong() for _ in range(5)]
        cgb, ex = C.c_double(), C.c_longlong()
        _check(lib.ausaxs_server_batch_stats(self._s, *map(C.byref, b), C.byref(cgb), C.byref(ex)))
        return dict(submitted=v[0].value, completed=v[1].value, failed=v[2].value, in_flight=fl.value,
                    peak_in_flight=pk.value, peak_gb=gb.value, pairs=pairs.value, kernel_pairs=kp.value,
                    cache_hits=b[0].value, cache_misses=b[1].value, batches=b[2].value, batched=b[3].value,
                    evictions=b[4].value, cached_gb=cgb.value, express=ex.value)


def pool_stats():
    """The caching allocator's counters (device bytes held / in use, hipMalloc and hipHostMalloc calls, reuses)."""
    held, used = C.c_size_t(), C.c_size_t()
    dm, hm, ru = C.c_longlong(), C.c_longlong(), C.c_longlong()
    _check(lib.ausaxs_pool_stats(C.byref(held), C.byref(used), C.byref(dm), C.byref(hm), C.byref(ru)))
    return dict(device_held=held.value, device_in_use=used.value, device_mallocs=dm.value, host_mallocs=hm.value,
                reuses=ru.value)


def pool_trim():
    _check(lib.ausaxs_pool_trim())


def cubic_spline(x, y, z):
    """math::CubicSpline: the natural cubic spline through (x, y) at z (host only)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    out = np.zeros(len(z), np.float64)
    _check(lib.ausaxs_cubic_spline(_ptr(x), _ptr(y), len(x), _ptr(z), len(z), _ptr(out)))
    return out


def lsq(x, y, err=None):
    """FP64 fits y ~ a x + b on the device for one model curve x[n] or many
    x[m][n]: rows of (a, b, var a, var b, chi2)."""
    x = np.ascontiguousarray(x, np.float64)
    single = x.ndim == 1
    x = x.reshape(1, -1) if single else x
    y = np.ascontiguousarray(y, np.float64)
    e = None if err is None else np.ascontiguousarray(err, np.float64)
    out = np.zeros((len(x), 5))
    _check(lib.ausaxs_lsq(_ptr(x), _ptr(y), _ptr(e), x.shape[1], len(x), _ptr(out)))
    return out[0] if single else out


def exact_debye(xyzw, q):
    """hist::exact_debye_transform on the device (XYZW points)."""
    p = _pts(xyzw)
    q = np.ascontiguousarray(q, np.float64)
    I = np.zeros(len(q))
    _check(lib.ausaxs_exact_debye(_ptr(p), len(p), _ptr(q), len(q), _ptr(I)))
    return I
