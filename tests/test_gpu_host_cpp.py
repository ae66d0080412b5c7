"""The C++ host mirror (basements_amd/host/mpm_solver.hpp) running the
reference's MPM unit tests, restated in tests/cpp/host_api_test.cpp; the
driver is built with the library (basements_amd.build)."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "basements_amd", "lib",
                   "host_api_test")


def test_cpp_host_mirror_runs_the_reference_unit_tests():
    if not os.path.exists(BIN):
        pytest.skip("host_api_test is not built")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith(("PASS", "FAIL"))]
    assert len(lines) == 9, r.stdout + r.stderr
    assert r.returncode == 0, r.stdout + r.stderr
