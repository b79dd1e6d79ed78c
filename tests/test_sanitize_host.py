"""The host-only modules (grid builder, radial hydration, synthetic structures,
threaded range split) built with g++ under AddressSanitizer +
UndefinedBehaviorSanitizer and under ThreadSanitizer, and run on the CPU
(tests/sanitize_host.cpp holds the checks). Any sanitizer report fails."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "ausaxs_amd" / "csrc"
SOURCES = [ROOT / "tests" / "sanitize_host.cpp", CSRC / "grid.cpp", CSRC / "hydrate.cpp", CSRC / "synth.cpp"]


@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_host_modules_under_sanitizer(tmp_path, sanitizer):
    exe = tmp_path / "sanitize_host"
    cmd = ["g++", "-std=c++20", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", str(CSRC), *map(str, SOURCES), "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
           "TSAN_OPTIONS": "halt_on_error=1", "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Sanitizer" not in r.stderr, r.stderr
    assert "all checks passed" in r.stdout
