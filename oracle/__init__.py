"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker. Imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg — never by the
product (ausaxs_amd).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
FFAVG, FFEXPLICIT = 0, 1
MAX_FF = 15

ENTRY = np.dtype([("value", np.float64), ("count", np.uint32), ("bin_center", np.float64)], align=True)
assert ENTRY.itemsize == 24


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_SIGS = {}


def _sig(name, restype, argtypes):
    """Declare the C signature of an oracle function (applied to every build loaded)."""
    _SIGS[name] = (restype, argtypes)
    f = getattr(lib, name)
    f.restype, f.argtypes = restype, argtypes


def _load(path=LIB_PATH):
    if not Path(path).exists():
        build()
    lib = C.CDLL(str(path))
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "orc_bin_pair": (None, [vp, vp, C.c_float, i, vp, vp]),
        "orc_rule_self": (i, [i, i, i]),
        "orc_rule_cross": (i, [i, i]),
        "orc_hist_ff": (i, [i, vp, i, vp, i, C.c_float, i, d, i, vp, vp, vp, vp]),
        "orc_pair_hist": (i, [vp, i, vp, i, C.c_float, i, i, vp]),
        "orc_hist_simple": (i, [vp, i, vp, i, C.c_float, i, i, vp, vp, vp, vp]),
        "orc_weighted_axis": (None, [vp, i, d, vp]),
        "orc_sinc_table": (None, [vp, i, vp, i, vp]),
        "orc_contract": (None, [vp, i, i, vp, i, vp]),
        "orc_contract_ld": (None, [vp, i, i, vp, i, vp]),
        "orc_ff_atomic": (d, [i, d]),
        "orc_ff_exv": (d, [i, d]),
        "orc_total": (None, [vp, i, vp]),
        "orc_exact_debye": (None, [vp, i, vp, i, vp]),
        "orc_lsq": (None, [vp, vp, vp, i, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args

    class Params(C.Structure):
        _fields_ = [(k, d) for k in ("cw", "cx", "crho", "dw_atomic", "dw_exv", "grid_v")]

    lib.orc_profiles.restype = None
    lib.orc_profiles.argtypes = [i, vp, i, vp, vp, vp, Params, vp]
    lib.Params = Params
    return lib


lib = _load()


@contextlib.contextmanager
def release_build():
    """Within the block every function here runs on the same restatement
    built with the reference's release flags for this host (Makefile target
    `release`): the CPU baseline's timing build. It is compiled into a
    temporary directory on the machine that times it (-march=native)."""
    global lib
    with tempfile.TemporaryDirectory() as tmp:
        out = Path(tmp) / "liboracle_release.so"
        subprocess.run(["make", "-s", "-C", str(HERE), "release", f"OUT={out}"], check=True)
        saved, lib = lib, _load(out)
        try:
            yield
        finally:
            lib = saved


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def inv_width(bin_width: float) -> np.float32:
    return np.float32(1.0 / bin_width)


def bin_pair(a, b, bin_width, rule):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    d = np.zeros(1, np.float32)
    k = np.zeros(1, np.int32)
    lib.orc_bin_pair(_p(a), _p(b), C.c_float(inv_width(bin_width)), rule, _p(d), _p(k))
    return float(d[0]), int(k[0])


def hist_ff(mode, atoms, waters, bin_width=0.25, bin_count=8000, z_exv=1.0, nthreads=1):
    """HistogramManagerMTFF{Avg,Explicit}::calculate_all -> dict of entry arrays (full bin_count) + max_bin."""
    atoms = np.ascontiguousarray(atoms, np.float32)
    waters = np.ascontiguousarray(waters if waters is not None else np.zeros((0, 4), np.float32), np.float32)
    aa = np.zeros(MAX_FF * MAX_FF * bin_count, ENTRY)
    aw = np.zeros(MAX_FF * bin_count, ENTRY)
    ww = np.zeros(bin_count, ENTRY)
    pt = np.zeros(bin_count, ENTRY)
    mb = lib.orc_hist_ff(mode, _p(atoms), len(atoms), _p(waters), len(waters), C.c_float(inv_width(bin_width)),
                         bin_count, z_exv, nthreads, _p(aa), _p(aw), _p(ww), _p(pt))
    if mb < 0:
        raise RuntimeError(f"oracle status {mb}")
    return dict(
        max_bin=mb,
        aa=aa.reshape(MAX_FF, MAX_FF, bin_count)[:, :, :mb],
        aw=aw.reshape(MAX_FF, bin_count)[:, :mb],
        ww=ww[:mb],
        p_tot=pt[:mb],
    )


def hist_simple(atoms, waters, bin_width=0.25, bin_count=8000, nthreads=1):
    atoms = np.ascontiguousarray(atoms, np.float32)
    waters = np.ascontiguousarray(waters if waters is not None else np.zeros((0, 4), np.float32), np.float32)
    out = [np.zeros(bin_count, ENTRY) for _ in range(4)]
    mb = lib.orc_hist_simple(_p(atoms), len(atoms), _p(waters), len(waters), C.c_float(inv_width(bin_width)),
                             bin_count, nthreads, *map(_p, out))
    if mb < 0:
        raise RuntimeError(f"oracle status {mb}")
    return dict(max_bin=mb, aa=out[0][:mb], aw=out[1][:mb], ww=out[2][:mb], p_tot=out[3][:mb])


def pair_hist(a, b=None, bin_width=0.25, bin_count=8000, factor=1):
    a = np.ascontiguousarray(a, np.float32)
    bb = None if b is None else np.ascontiguousarray(b, np.float32)
    out = np.zeros(bin_count, ENTRY)
    rc = lib.orc_pair_hist(_p(a), len(a), _p(bb), 0 if bb is None else len(bb), C.c_float(inv_width(bin_width)),
                           bin_count, factor, _p(out))
    if rc:
        raise RuntimeError(f"oracle status {rc}")
    return out


def weighted_axis(p, bin_width):
    p = np.ascontiguousarray(p, ENTRY)
    d = np.zeros(len(p), np.float64)
    lib.orc_weighted_axis(_p(p), len(p), bin_width, _p(d))
    return d


def sinc_table(q, d):
    q = np.ascontiguousarray(q, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    out = np.zeros((len(q), len(d)), np.float64)
    lib.orc_sinc_table(_p(q), len(q), _p(d), len(d), _p(out))
    return out


def contract(H, S, long_double=False):
    H = np.ascontiguousarray(H, np.float64)
    S = np.ascontiguousarray(S, np.float64)
    out = np.zeros((H.shape[0], S.shape[0]), np.float64)
    f = lib.orc_contract_ld if long_double else lib.orc_contract
    f(_p(H), H.shape[0], H.shape[1], _p(S), S.shape[0], _p(out))
    return out


def ff_atomic(t, q):
    return lib.orc_ff_atomic(t, q)


def ff_exv(t, q):
    return lib.orc_ff_exv(t, q)


def profiles(mode, q, sinqd_aa, sinqd_aw, sinqd_ww, cw=1.0, cx=1.0, crho=1.0, dw_atomic=0.0, dw_exv=0.0,
             grid_exv_volume=1.0):
    q = np.ascontiguousarray(q, np.float64)
    nq = len(q)
    a3 = np.ascontiguousarray(sinqd_aa, np.float64).reshape(MAX_FF * MAX_FF * nq)
    a2 = np.ascontiguousarray(sinqd_aw, np.float64).reshape(MAX_FF * nq)
    a1 = np.ascontiguousarray(sinqd_ww, np.float64).reshape(nq)
    out = np.zeros((6, nq), np.float64)
    lib.orc_profiles(mode, _p(q), nq, _p(a3), _p(a2), _p(a1),
                     lib.Params(cw, cx, crho, dw_atomic, dw_exv, grid_exv_volume), _p(out))
    return out


def total(prof):
    prof = np.ascontiguousarray(prof, np.float64)
    nq = prof.shape[1]
    out = np.zeros(nq, np.float64)
    lib.orc_total(_p(prof), nq, _p(out))
    return out


def exact_debye(xyzw, q):
    xyzw = np.ascontiguousarray(xyzw, np.float32)
    q = np.ascontiguousarray(q, np.float64)
    out = np.zeros(len(q), np.float64)
    lib.orc_exact_debye(_p(xyzw), len(xyzw), _p(q), len(q), _p(out))
    return out


def lsq(x, y, err=None):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    e = None if err is None else np.ascontiguousarray(err, np.float64)
    out = np.zeros(5, np.float64)
    lib.orc_lsq(_p(x), _p(y), _p(e), len(x), _p(out))
    return out


# ---------------------------------------------------------------------------
# full FF pipeline: histograms -> weighted axis -> sinc -> sinqd -> profiles
# ---------------------------------------------------------------------------
def debye_ff(mode, hist, q, bin_width, weighted=True, long_double=False, **params):
    """I(q) and the six profiles from an oracle FF histogram (FP64 throughout)."""
    mb = hist["max_bin"]
    if weighted:
        d = weighted_axis(hist["p_tot"], bin_width)
    else:
        d = np.arange(mb) * ((mb * bin_width) / mb)
        d[0] = 0
    S = sinc_table(q, d)
    nq = len(q)
    aa = contract(hist["aa"]["value"].reshape(-1, mb), S, long_double).reshape(MAX_FF, MAX_FF, nq)
    aw = contract(hist["aw"]["value"].reshape(-1, mb), S, long_double).reshape(MAX_FF, nq)
    ww = contract(hist["ww"]["value"].reshape(1, mb), S, long_double).reshape(nq)
    prof = profiles(mode, q, aa, aw, ww, **params)
    return total(prof), prof


# ---------------------------------------------------------------------------
# grid excluded volume (FFGrid)
# ---------------------------------------------------------------------------
_sig("orc_grid_exv", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                             C.c_void_p])
_sig("orc_hist_ffgrid", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double,
                                C.c_float, C.c_int, C.c_int] + [C.c_void_p] * 6)
_sig("orc_weighted_axis_n", None, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p])
FFGRID = 2


def grid_exv(atoms_xyz, ff, waters_xyz=None):
    """Grid dummies (n x 3 FP64, reference order) and grid volume."""
    a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ff, np.int32).reshape(-1)
    w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)), np.float64).reshape(-1, 3)
    cap = max(1, len(a) * 64)
    out = np.zeros((cap, 3), np.float64)
    vol = C.c_double()
    n = lib.orc_grid_exv(_p(a), _p(f), len(a), _p(w), len(w), _p(out), cap, C.byref(vol))
    if n < 0:
        raise RuntimeError("grid_exv capacity")
    return out[:n].copy(), vol.value


def pack_ff(xyz, ff):
    out = np.zeros((len(xyz), 4), np.float32)
    out[:, :3] = np.asarray(xyz, np.float64).astype(np.float32)
    out.view(np.int32)[:, 3] = ff
    return out


def hist_ffgrid(atoms_xyz, ff, waters_xyz=None, bin_width=0.25, bin_count=8000, nthreads=1, exv_scale=1.0):
    """HistogramManagerMTFFGrid::calculate_all from FP64 coordinates; exv_scale != 1:
    HistogramManagerMTFFGridScalableExv's eval_scaled_exv (float dummy coordinates times k)."""
    w = np.zeros((0, 3)) if waters_xyz is None else np.asarray(waters_xyz, np.float64)
    dummies, vol = grid_exv(atoms_xyz, ff, w)
    atoms = pack_ff(atoms_xyz, ff)
    waters = pack_ff(w, np.ones(len(w), np.int32))
    xs = pack_ff(dummies, np.zeros(len(dummies), np.int32))
    if exv_scale != 1.0:
        xs[:, :3] = (xs[:, :3].astype(np.float64) * exv_scale).astype(np.float32)
    nb = bin_count
    aa = np.zeros(MAX_FF * MAX_FF * nb, ENTRY)
    aw = np.zeros(MAX_FF * nb, ENTRY)
    ww, pt, pax, pxx = (np.zeros(nb, ENTRY) for _ in range(4))
    mb = lib.orc_hist_ffgrid(_p(atoms), len(atoms), _p(waters), len(waters), _p(xs), len(xs), vol,
                             C.c_float(inv_width(bin_width)), nb, nthreads, _p(aa), _p(aw), _p(ww), _p(pt), _p(pax),
                             _p(pxx))
    if mb < 0:
        raise RuntimeError(f"oracle status {mb}")
    d_ax = np.zeros(mb)
    d_xx = np.zeros(mb)
    lib.orc_weighted_axis_n(_p(pax), mb, nb, bin_width, _p(d_ax))
    lib.orc_weighted_axis_n(_p(pxx), mb, nb, bin_width, _p(d_xx))
    return dict(max_bin=mb, aa=aa.reshape(MAX_FF, MAX_FF, nb)[:, :, :mb], aw=aw.reshape(MAX_FF, nb)[:, :mb],
                ww=ww[:mb], p_tot=pt[:mb], p_tot_ax=pax[:mb], p_xx=pxx[:mb], d_ax=d_ax, d_xx=d_xx, volume=vol,
                dummies=dummies)


def debye_ffgrid(hist, q, bin_width, long_double=False, **params):
    """I(q) and profiles of an oracle FFGrid histogram (CompositeDistanceHistogramFFGrid.cpp:77-119)."""
    mb = hist["max_bin"]
    d = weighted_axis(hist["p_tot"], bin_width)  # count 0: the centres themselves, d[0] = 0
    S_aa, S_ax, S_xx = sinc_table(q, d), sinc_table(q, hist["d_ax"]), sinc_table(q, hist["d_xx"])
    nq = len(q)
    aa_v = hist["aa"]["value"]
    sa = contract(aa_v.reshape(-1, mb), S_aa, long_double).reshape(MAX_FF, MAX_FF, nq)
    sx = contract(aa_v[:, 0, :], S_ax, long_double)
    sa[:, 0, :] = sx
    sa[0, 0, :] = contract(aa_v[0, 0][None, :], S_xx, long_double)[0]
    aw_v = hist["aw"]["value"]
    sw = contract(aw_v, S_aa, long_double)
    sw[0] = contract(aw_v[0][None, :], S_ax, long_double)[0]
    s_ww = contract(hist["ww"]["value"][None, :], S_aa, long_double)[0]
    prof = profiles(FFGRID, q, sa, sw, s_ww, **params)
    return total(prof), prof


# ---------------------------------------------------------------------------
# grid excluded volume with a surface layer (FFGridSurface)
# ---------------------------------------------------------------------------
_sig("orc_grid_exv_surface", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double])
_sig("orc_hist_ffgrid_surface", C.c_int, ([C.c_void_p, C.c_int] * 4 + [C.c_double, C.c_float, C.c_int, C.c_int]
                                        + [C.c_void_p] * 13))


def grid_exv_surface(atoms_xyz, ff, waters_xyz=None, rvol=2.15, atomic_radius=None, surface_thickness=1.0):
    """Interior and surface dummies (each n x 3 FP64, reference order) and the grid volume.
    rvol: the minimum excluded-volume radius; atomic_radius: one radius for every atom in
    place of the van der Waals radii; surface_thickness in A (cell width 1)."""
    a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ff, np.int32).reshape(-1)
    w = np.ascontiguousarray(waters_xyz if waters_xyz is not None else np.zeros((0, 3)), np.float64).reshape(-1, 3)
    cap = max(1, len(a) * (2 * int(np.ceil(max(rvol, 2.15))) + 1) ** 3)
    xi = np.zeros((cap, 3), np.float64)
    xs = np.zeros((cap, 3), np.float64)
    ns, vol = C.c_int(), C.c_double()
    n = lib.orc_grid_exv_surface(_p(a), _p(f), len(a), _p(w), len(w), _p(xi), _p(xs), cap, C.byref(ns),
                                 C.byref(vol), rvol, -1.0 if atomic_radius is None else atomic_radius,
                                 float(surface_thickness))
    if n < 0:
        raise RuntimeError("grid_exv_surface capacity")
    return xi[:n].copy(), xs[:ns.value].copy(), vol.value


_sig("orc_hydrate_radial", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_void_p])


class HydrateOpts(C.Structure):
    _fields_ = [("atomic_radius", C.c_double), ("hydration_radius", C.c_double), ("shell_correction", C.c_double),
                ("explicit_grid", C.c_int), ("grid_lim", C.c_double * 6)]


_sig("orc_hydrate_radial_opts", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                          C.POINTER(HydrateOpts), C.c_void_p, C.c_int, C.c_void_p])


def hydrate_radial(atoms_xyz, ff, bodies=None, noise=None, atomic_radius=-1.0, hydration_radius=1.71,
                   shell_correction=-0.35, grid_limits=None):
    """Radial hydration shell (see orc_hydrate_radial_opts): waters (m x 3) and
    per-body counts. grid_limits: ((xmin, ymin, zmin), (xmax, ymax, zmax)) for
    Grid(Limit3D), None for the grid built from the atoms."""
    a = np.ascontiguousarray(atoms_xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ff, np.int32).reshape(-1)
    b = np.ascontiguousarray([len(a)] if bodies is None else bodies, np.int32)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64).reshape(len(a), 26, 3)
    cap = max(1, 26 * len(a))
    out = np.zeros((cap, 3), np.float64)
    per = np.zeros(len(b), np.int32)
    o = HydrateOpts(atomic_radius, hydration_radius, shell_correction, 0)
    if grid_limits is not None:
        o.explicit_grid = 1
        for k in range(3):
            o.grid_lim[k], o.grid_lim[3 + k] = grid_limits[0][k], grid_limits[1][k]
    m = lib.orc_hydrate_radial_opts(_p(a), _p(f), len(a), _p(b), len(b), _p(nz), o, _p(out), cap, _p(per))
    if m == -2:
        raise ValueError("an atom lies outside the grid")
    if m < 0:
        raise RuntimeError("hydrate_radial capacity")
    return out[:m].copy(), per


def hist_ffgrid_surface(atoms_xyz, ff, waters_xyz=None, bin_width=0.25, bin_count=8000, nthreads=1,
                        surface_thickness=1.0):
    """HistogramManagerMTFFGridSurface::calculate_all from FP64 coordinates (see orc_hist_ffgrid_surface)."""
    w = np.zeros((0, 3)) if waters_xyz is None else np.asarray(waters_xyz, np.float64)
    xi_xyz, xs_xyz, vol = grid_exv_surface(atoms_xyz, ff, w, surface_thickness=surface_thickness)
    atoms = pack_ff(atoms_xyz, ff)
    waters = pack_ff(w, np.ones(len(w), np.int32))
    xi = pack_ff(xi_xyz, np.zeros(len(xi_xyz), np.int32))
    xs = pack_ff(xs_xyz, np.zeros(len(xs_xyz), np.int32))
    nb = bin_count
    aa = np.zeros(MAX_FF * MAX_FF * nb, ENTRY)
    aw, axi, axs = (np.zeros(MAX_FF * nb, ENTRY) for _ in range(3))
    ww, pt, pax, pxx, xxi, xxs, xxc, wxi, wxs = (np.zeros(nb, ENTRY) for _ in range(9))
    mb = lib.orc_hist_ffgrid_surface(_p(atoms), len(atoms), _p(waters), len(waters), _p(xi), len(xi), _p(xs),
                                     len(xs), vol, C.c_float(inv_width(bin_width)), nb, nthreads, _p(aa), _p(aw),
                                     _p(ww), _p(pt), _p(pax), _p(pxx), _p(xxi), _p(xxs), _p(xxc), _p(wxi), _p(wxs),
                                     _p(axi), _p(axs))
    if mb < 0:
        raise RuntimeError(f"oracle status {mb}")
    d_ax = np.zeros(mb)
    d_xx = np.zeros(mb)
    lib.orc_weighted_axis_n(_p(pax), mb, nb, bin_width, _p(d_ax))
    lib.orc_weighted_axis_n(_p(pxx), mb, nb, bin_width, _p(d_xx))
    return dict(max_bin=mb, aa=aa.reshape(MAX_FF, MAX_FF, nb)[:, :, :mb], aw=aw.reshape(MAX_FF, nb)[:, :mb],
                ww=ww[:mb], p_tot=pt[:mb], p_tot_ax=pax[:mb], p_xx=pxx[:mb], d_ax=d_ax, d_xx=d_xx, volume=vol,
                xx_interior=xxi[:mb], xx_surface=xxs[:mb], xx_cross=xxc[:mb], wx_interior=wxi[:mb],
                wx_surface=wxs[:mb], ax_interior=axi.reshape(MAX_FF, nb)[:, :mb],
                ax_surface=axs.reshape(MAX_FF, nb)[:, :mb], interior=xi_xyz, surface=xs_xyz)


def exv_factor(q, cx):
    """CompositeDistanceHistogramFFGridSurface::exv_factor (.cpp:63-66), r_m = 1.62 A."""
    rm2 = 1.62 * 1.62 / 4
    return cx ** 3 * np.exp(-rm2 * (cx ** 2 - 1) * np.asarray(q, np.float64) ** 2)


def debye_ffgrid_surface(hist, q, bin_width, cw=1.0, cx=1.0, crho=1.0, grid_exv_volume=1.0):
    """I(q) and the six profiles {aa, ax, aw, xx, wx, ww} of an oracle FFGridSurface histogram
    (CompositeDistanceHistogramFFGridSurface::cache_refresh_intensity_profiles, .cpp:174-266):
    the surface profiles scale by exv_factor(q, cx) per q; atomic terms on the p_tot axis, the
    ax / wx sums on the p_tot_ax axis and xx on the p_xx axis. No Debye-Waller factors."""
    q = np.asarray(q, np.float64)
    nq = len(q)
    mb = hist["max_bin"]
    S_aa = sinc_table(q, weighted_axis(hist["p_tot"], bin_width))
    S_ax, S_xx = sinc_table(q, hist["d_ax"]), sinc_table(q, hist["d_xx"])
    fa = np.array([[ff_atomic(f, qk) for f in range(MAX_FF)] for qk in q])  # nq x MAX_FF
    V = grid_exv_volume
    fx = 0.334 * V * np.exp(-(V ** (2 / 3) / (4 * np.pi)) * q * q)  # ExvFormFactor.h:23-34
    fw = fa[:, 1]
    c = exv_factor(q, cx)
    prof = np.zeros((6, nq))
    aa_v = hist["aa"]["value"]
    sa = contract(aa_v[1:, 1:].reshape(-1, mb), S_aa).reshape(MAX_FF - 1, MAX_FF - 1, nq)
    prof[0] = np.einsum("abq,qa,qb->q", sa, fa[:, 1:], fa[:, 1:])
    si = contract(hist["ax_interior"]["value"][1:], S_ax)  # (MAX_FF - 1) x nq
    ss = contract(hist["ax_surface"]["value"][1:], S_ax)
    prof[1] = 2 * crho * np.sum((si + c[None, :] * ss) * fa[:, 1:].T, axis=0) * fx
    xi, xs, xc = (contract(hist[k]["value"][None, :], S_xx)[0] for k in ("xx_interior", "xx_surface", "xx_cross"))
    prof[3] = crho * crho * (xi + c * c * xs + c * xc) * fx * fx
    sw = contract(hist["aw"]["value"][1:], S_aa)
    prof[2] = 2 * cw * np.sum(sw * fa[:, 1:].T, axis=0) * fw
    wi, ws = (contract(hist[k]["value"][None, :], S_ax)[0] for k in ("wx_interior", "wx_surface"))
    prof[4] = 2 * crho * cw * (wi + c * ws) * fw * fx
    prof[5] = cw * cw * contract(hist["ww"]["value"][None, :], S_aa)[0] * fw * fw
    return total(prof), prof


# ---------------------------------------------------------------------------
# rigid bodies (PartialHistogramManagerMT) and the plain composite I(q)
# ---------------------------------------------------------------------------
_sig("orc_rigid_transform", None, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p])
_sig("orc_hist_partial", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float,
                                 C.c_int, C.c_int] + [C.c_void_p] * 4)


def rigid_transform(xyz, w, T):
    """FP64 rigid move then float cast (n x 3 doubles, n weights, T[12]) -> n x 4 XYZW floats."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    w = np.ascontiguousarray(w, np.float32)
    T = np.ascontiguousarray(T, np.float64)
    out = np.zeros((len(xyz), 4), np.float32)
    lib.orc_rigid_transform(_p(xyz), _p(w), len(xyz), _p(T), _p(out))
    return out


def hist_partial(self_bodies, bodies, sizes, waters=None, bin_width=0.25, bin_count=8000, nthreads=1):
    """PartialHistogramManagerMT::calculate_all after every body moved (see orc_hist_partial)."""
    self_bodies = np.ascontiguousarray(self_bodies, np.float32)
    bodies = np.ascontiguousarray(bodies, np.float32)
    sizes = np.ascontiguousarray(sizes, np.int32)
    waters = np.ascontiguousarray(waters if waters is not None else np.zeros((0, 4), np.float32), np.float32)
    out = [np.zeros(bin_count, ENTRY) for _ in range(4)]
    mb = lib.orc_hist_partial(_p(self_bodies), _p(bodies), _p(sizes), len(sizes), _p(waters), len(waters),
                              C.c_float(inv_width(bin_width)), bin_count, nthreads, *map(_p, out))
    if mb < 0:
        raise RuntimeError(f"oracle status {mb}")
    return dict(max_bin=mb, aa=out[0][:mb], aw=out[1][:mb], ww=out[2][:mb], p_tot=out[3][:mb])


_sig("orc_affine_float", None, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p])
_sig("orc_hist_symmetry", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int, C.c_int] + [C.c_void_p] * 4)


def affine_float(pts, T):
    """A symmetry copy of float XYZW points (see orc_affine_float)."""
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 4)
    T = np.ascontiguousarray(T, np.float64).reshape(12)
    out = np.zeros_like(pts)
    lib.orc_affine_float(_p(pts), len(pts), _p(T), _p(out))
    return out


_sig("orc_hist_symmetry_partial", C.c_int, [C.c_void_p] * 3 + [C.c_void_p, C.c_int] + [C.c_void_p] * 6 + [
    C.c_int, C.c_float, C.c_int, C.c_int] + [C.c_void_p] * 4)


def hist_symmetry(bodies, symmetries, waters=None, bin_width=0.25, bin_count=8000, nthreads=1, partial=False,
                  self_bodies=None, self_scale=None):
    """SymmetryManagerMT::calculate_all (see orc_hist_symmetry), or with
    partial=True PartialSymmetryManagerMT's (orc_hist_symmetry_partial; self
    terms from self_bodies with factors self_scale, default: the bodies and
    1 + their copies). bodies: list of (n, 4) float arrays; symmetries: per
    body a list of (transforms (reps, 12), schedule (k, 3))."""
    atoms = np.ascontiguousarray(np.concatenate(bodies) if bodies else np.zeros((0, 4)), np.float32)
    sizes = np.array([len(b) for b in bodies], np.int32)
    per = np.array([len(x) for x in symmetries], np.int32)
    flat = [x for xs in symmetries for x in xs]
    reps = np.array([len(np.asarray(T).reshape(-1, 12)) for T, _ in flat], np.int32)
    T = np.ascontiguousarray(np.concatenate([np.asarray(T, np.float64).reshape(-1, 12) for T, _ in flat])
                             if flat else np.zeros((0, 12)), np.float64)
    slen = np.array([len(np.asarray(S).reshape(-1, 3)) for _, S in flat], np.int32)
    sched = np.ascontiguousarray(np.concatenate([np.asarray(S, np.int32).reshape(-1, 3) for _, S in flat])
                                 if flat else np.zeros((0, 3)), np.int32)
    waters = np.ascontiguousarray(waters if waters is not None else np.zeros((0, 4), np.float32), np.float32)
    out = [np.zeros(bin_count, ENTRY) for _ in range(4)]
    if partial:
        sb = np.ascontiguousarray(np.concatenate(self_bodies) if self_bodies is not None else atoms, np.float32)
        if self_scale is None:
            self_scale = [1 + sum(len(np.asarray(T).reshape(-1, 12)) for T, _ in xs) for xs in symmetries]
        sc = np.ascontiguousarray(self_scale, np.int32)
        mb = lib.orc_hist_symmetry_partial(_p(atoms), _p(sb), _p(sc), _p(sizes), len(sizes), _p(per), _p(reps),
                                           _p(T), _p(slen), _p(sched), _p(waters), len(waters),
                                           C.c_float(inv_width(bin_width)), bin_count, nthreads, *map(_p, out))
    else:
        mb = lib.orc_hist_symmetry(_p(atoms), _p(sizes), len(sizes), _p(per), _p(reps), _p(T), _p(slen), _p(sched),
                                   _p(waters), len(waters), C.c_float(inv_width(bin_width)), bin_count, nthreads,
                                   *map(_p, out))
    if mb < 0:
        raise RuntimeError(f"oracle status {mb}")
    return dict(max_bin=mb, aa=out[0][:mb], aw=out[1][:mb], ww=out[2][:mb], p_tot=out[3][:mb])


_sig("orc_spline", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p])


def spline(x, y, z):
    """math::CubicSpline(x, y).spline(z) (see orc_spline)."""
    x, y = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(np.atleast_1d(z), np.float64)
    out = np.zeros_like(z)
    if lib.orc_spline(_p(x), _p(y), len(x), _p(z), len(z), _p(out)):
        raise ValueError("a spline needs at least four knots")
    return out


def golden(f, lo, hi, tol=1e-4):
    """mini::Golden::search + minimize_override (Golden.cpp:22-70): the
    centre of the final interval, its half width and the evaluation count."""
    invphi = 1 / ((1 + np.sqrt(5.0)) / 2)
    invphi2 = invphi * invphi
    evals = [0]

    def F(v):
        evals[0] += 1
        return f(v)
    a, b = min(lo, hi), (lo + hi) - min(lo, hi)
    diff = b - a
    if diff >= tol:
        n = int(np.ceil(np.log(tol / diff) / np.log(invphi)))
        c, d = a + invphi2 * diff, a + invphi * diff
        fc, fd = F(c), F(d)
        for _ in range(n - 1):
            if fc < fd:
                b, d, fd = d, c, fc
                diff = invphi * diff
                c = a + invphi2 * diff
                fc = F(c)
            else:
                a, c, fc = c, d, fd
                diff = invphi * diff
                d = a + invphi * diff
                fd = F(d)
        a, b = (a, d) if fc < fd else (c, b)
    centre = a + (b - a) / 2
    F(centre)
    return centre, (b - a) / 2, evals[0]


def debye_simple(hist, q, bin_width, cw=1.0):
    """CompositeDistanceHistogram::debye_transform (CompositeDistanceHistogram.cpp:55-70 with
    DistanceHistogram.cpp:54-87): p = aa + cw*aw + cw^2*ww on p_tot's weighted axis, times exp(-q^2)."""
    d = weighted_axis(hist["p_tot"], bin_width)
    p = hist["aa"]["value"] + cw * hist["aw"]["value"] + cw * cw * hist["ww"]["value"]
    q = np.asarray(q, np.float64)
    return contract(p[None, :], sinc_table(q, d))[0] * np.exp(-q * q)
