"""The host-side particle transposes (basements_amd/csrc/host_transpose.hpp:
the upload transpose with its finiteness check, the permuted download and
the order-range check, all on up to 16 host threads) under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer
(host CPU only): tests/cpp/staging_sanitize.cpp drives them at serial and
multi-threaded sizes with exact-size buffers; any sanitizer report or failed
check fails the run."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SAN = {
    "asan_ubsan": ("address,undefined", dict(ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
                                             UBSAN_OPTIONS="print_stacktrace=1")),
    "tsan": ("thread", dict(TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")),
}


@pytest.mark.parametrize("kind", sorted(SAN))
def test_host_transposes_are_clean_under_sanitizer(tmp_path, kind):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    flag, env_add = SAN[kind]
    exe = tmp_path / f"staging_{kind}"
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={flag}",
           "-fno-sanitize-recover=all", "-pthread", "-I", os.path.join(ROOT, "basements_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "staging_sanitize.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=dict(os.environ, **env_add), timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "clean" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
