"""C-ABI library checks that need no GPU: it loads, exports every declared
symbol, runs its host-only entry points, and fails loudly (no CPU fallback)
when a GPU entry point is called without a device."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ausaxs_amd as A

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "ausaxs_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ausaxs_\w+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ausaxs_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(A.lib, n)


def test_library_is_gfx950():
    # the embedded offload bundle names its target triple
    assert b"amdgcn-amd-amdhsa--gfx950" in A.LIB_PATH.read_bytes()


def test_default_settings_mirror_reference():
    s = A.default_settings()
    # HistogramSettings.cpp: width 0.25 (d_axis 0..2000 in 8000 bins), weighted bins on
    assert s.bin_width == 0.25 and s.bin_count == 8000 and s.weighted_bins == 1


def test_default_q_axis():
    q = A.default_q_axis(1e-4, 0.5)
    # ConstantsAxes.h:14 q_axis(1e-4, 1, 200); sub_axis(1e-4, 0.5) keeps 99 points
    assert len(q) == 99
    np.testing.assert_allclose(q, 1e-4 + np.arange(99) * (1 - 1e-4) / 200, rtol=0, atol=1e-15)
    assert len(A.default_q_axis(1e-4, 1.0)) == 200


def test_synth_deterministic_and_avoiding():
    a = A.synth_globule(1, 2000, (20, 20, 20), 2.0)
    b = A.synth_globule(1, 2000, (20, 20, 20), 2.0)
    assert np.array_equal(a, b)
    xyz = a[:, :3].astype(np.float64)
    assert np.all(np.sum(xyz ** 2, axis=1) <= 20.0 ** 2 * (1 + 1e-6))
    d2 = np.sum((xyz[:, None, :] - xyz[None, :, :]) ** 2, axis=-1)
    np.fill_diagonal(d2, np.inf)
    assert np.sqrt(d2.min()) >= 1.9  # exclusion 2.0 (relaxed by at most a few % when packing jams)
    ff = a.view(np.int32)[:, 3]
    frac = {t: np.mean(ff == t) for t in (3, 7, 11, 12)}
    assert abs(frac[3] - 0.62) < 0.04 and abs(frac[7] - 0.17) < 0.03 and abs(frac[11] - 0.20) < 0.03
    w = A.synth_shell(2, 500, (20, 20, 20), 3.0)
    r = np.linalg.norm(w[:, :3], axis=1)
    np.testing.assert_allclose(r, 23.0, rtol=1e-5)
    assert np.all(w.view(np.int32)[:, 3] == 1)


def test_invalid_inputs_are_errors():
    with pytest.raises(A.AusaxsError):
        A._check(A.lib.ausaxs_synth_globule(1, 10, -1.0, 1.0, 1.0, 2.0, None))


def _has_gpu():
    try:
        A.device_info()
        return True
    except A.AusaxsError:
        return False


def test_gpu_entry_fails_loudly_without_device():
    if _has_gpu():
        pytest.skip("a HIP device is present")
    atoms = A.points_ff(np.zeros((4, 3)), 3)
    with pytest.raises(A.AusaxsError, match="no HIP device"):
        A.calculate_all(A.MODE_FFEXPLICIT, atoms)


def _compile(cmd, cwd):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    assert r.returncode == 0, r.stdout + r.stderr


def test_header_is_plain_c(tmp_path):
    # the header compiles as strict C99 and as C++20 with every warning an error
    src = tmp_path / "h.c"
    src.write_text('#include "ausaxs_hip.h"\nint main(void) { ausaxs_settings s; ausaxs_default_settings(&s);'
                   " return s.bin_count == 8000 ? 0 : 1; }\n")
    inc = str(ROOT / "include")
    _compile(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", inc, "-fsyntax-only",
              str(src)], tmp_path)
    _compile(["g++", "-std=c++20", "-pedantic", "-Wall", "-Wextra", "-Werror", "-x", "c++", "-I", inc,
              "-fsyntax-only", str(src)], tmp_path)


def test_cpp_consumer_builds_against_header(tmp_path):
    # a caller that binds the histogram managers through the header alone
    # (tests/capi_consumer.cpp): it compiles warning-free, links, and on a
    # host without a GPU stops at the first compute call with the library's
    # message instead of computing anything on the CPU
    import json

    exe = tmp_path / "capi_consumer"
    _compile(["g++", "-std=c++20", "-O1", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"),
              str(ROOT / "tests" / "capi_consumer.cpp"), "-o", str(exe), "-L", str(A.LIB_PATH.parent),
              "-lausaxs_hip", f"-Wl,-rpath,{A.LIB_PATH.parent}"], tmp_path)
    if _has_gpu():
        pytest.skip("a HIP device is present (the GPU test runs the consumer)")
    r = subprocess.run([str(exe), "200"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "no HIP device" in json.loads(r.stdout)["error"]
