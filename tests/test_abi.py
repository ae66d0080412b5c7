"""The C ABI without a GPU: the library loads, exports every function the
header declares, its structs have the layout the Python mirror assumes, and
argument checks that precede any device call answer with the documented
status codes."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bm_mpm.h")


def _lib_or_skip():
    from basements_amd import lib
    from basements_amd.mpm import LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.skip("libbm_mpm.so is not built")
    return lib()


def test_library_exports_every_header_symbol():
    from basements_amd import header_symbols

    L = _lib_or_skip()
    names = header_symbols(HEADER)
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.bm_abi_version() == 1


def test_python_signatures_cover_the_header():
    from basements_amd import header_symbols
    from basements_amd.mpm import _SIGS

    assert sorted(_SIGS) == header_symbols(HEADER)


def test_argument_errors_without_device():
    from basements_amd.mpm import Desc

    L = _lib_or_skip()
    h = C.c_void_p()
    assert L.bm_mpm_create(None, C.byref(h)) == 1  # BM_ERR_INVALID_ARG
    assert b"null" in L.bm_last_error()
    d = Desc()
    d.dx = -1.0
    d.grid_dims[:] = [8, 8, 8]
    assert L.bm_mpm_create(C.byref(d), C.byref(h)) == 1
    assert b"dx" in L.bm_last_error()
    d.dx = 0.1
    d.mode = 5
    assert L.bm_mpm_create(C.byref(d), C.byref(h)) == 1
    assert L.bm_mpm_step(None, 1e-4) == 1
    # a rollout count outside 0..BM_DIFF_MAX_ROLLOUTS is refused before any device call
    from basements_amd.mpm import DiffDescC

    dd = DiffDescC()
    dd.dx = 0.01
    dd.grid_dims[:] = [8, 8, 8]
    dd.max_particles, dd.precision = 16, 32
    for bad in (-1, 4097):
        dd.rollouts = bad
        assert L.bm_diff_create(C.byref(dd), C.byref(h)) == 1
        assert b"rollouts" in L.bm_last_error()
    assert L.bm_mpm_destroy(None) == 0


_LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "bm_mpm.h"
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m))
#define S(T) printf(#T " sizeof %zu\n", sizeof(T))
int main(void) {
  S(bm_mpm_desc); F(bm_mpm_desc, dx); F(bm_mpm_desc, grid_origin); F(bm_mpm_desc, grid_dims);
  F(bm_mpm_desc, max_particles); F(bm_mpm_desc, mode); F(bm_mpm_desc, sort_every); F(bm_mpm_desc, stream);
  S(bm_particles); F(bm_particles, V0);
  S(bm_particles_device); F(bm_particles_device, capacity);
  S(bm_collider); F(bm_collider, quat); F(bm_collider, friction); F(bm_collider, shape);
  S(bm_reaction); F(bm_reaction, nodes_inside);
  S(bm_body); F(bm_body, inv_mass); F(bm_body, integrate); F(bm_body, applied_torque); F(bm_body, shape);
  S(bm_batch_desc); F(bm_batch_desc, max_scenes); F(bm_batch_desc, stream);
  S(bm_scene_desc); F(bm_scene_desc, young_modulus); F(bm_scene_desc, gravity); F(bm_scene_desc, coupler);
  S(bm_server_desc); F(bm_server_desc, hbm_budget); F(bm_server_desc, dt); F(bm_server_desc, max_bypass);
  S(bm_request); F(bm_request, n); F(bm_request, output); F(bm_request, substeps); F(bm_request, arrival);
  S(bm_request_info); F(bm_request_info, arrival); F(bm_request_info, body); F(bm_request_info, reaction);
  S(bm_diff_desc); F(bm_diff_desc, max_particles); F(bm_diff_desc, reference_svd); F(bm_diff_desc, rollouts);
  F(bm_diff_desc, stream);
  S(bm_diff_params); F(bm_diff_params, has_tool); F(bm_diff_params, tool_center); F(bm_diff_params, tool_friction);
  S(bm_diff_grad); F(bm_diff_grad, young_modulus); F(bm_diff_grad, lambda);
  return 0;
}
"""


def test_struct_layouts_match_the_python_mirror(tmp_path):
    from basements_amd import mpm

    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text(_LAYOUT_C)
    exe = tmp_path / "layout"
    inc = os.path.join(ROOT, "include")
    r = subprocess.run([cc, "-std=c11", f"-I{inc}", str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    want = {}
    for line in out:
        if line:
            t, m, v = line.split()
            want[(t, m)] = int(v)
    py = {"bm_mpm_desc": mpm.Desc, "bm_particles": mpm.HostParticles, "bm_particles_device": mpm.DeviceParticles,
          "bm_collider": mpm.ColliderC, "bm_reaction": mpm.ReactionC, "bm_body": mpm.BodyC,
          "bm_batch_desc": mpm.BatchDescC, "bm_scene_desc": mpm.SceneDescC, "bm_server_desc": mpm.ServerDescC,
          "bm_request": mpm.RequestC, "bm_request_info": mpm.RequestInfoC, "bm_diff_desc": mpm.DiffDescC,
          "bm_diff_params": mpm.DiffParamsC, "bm_diff_grad": mpm.DiffGradC}
    for (t, m), v in want.items():
        m = "lambda_" if m == "lambda" else m
        if m == "sizeof":
            assert C.sizeof(py[t]) == v, (t, C.sizeof(py[t]), v)
        else:
            assert getattr(py[t], m).offset == v, (t, m, getattr(py[t], m).offset, v)


def test_batch_footprint_is_linear():
    from basements_amd import batch_footprint

    _lib_or_skip()
    z = batch_footprint(0, 0, 0)
    p, b, s = batch_footprint(1000, 0, 0) - z, batch_footprint(0, 1000, 0) - z, batch_footprint(0, 0, 10) - z
    assert batch_footprint(1000, 1000, 10) == z + p + b + s
    # particle state alone is 2 x 26 floats; a block's two grids 2 x 64 x 24 bytes
    assert p >= 1000 * 208 and b >= 1000 * 3072 and s > 0


def test_server_argument_errors_without_device():
    from basements_amd.mpm import ServerDescC

    L = _lib_or_skip()
    h = C.c_void_p()
    assert L.bm_server_create(None, C.byref(h)) == 1
    d = ServerDescC()
    d.dt = 0.0
    assert L.bm_server_create(C.byref(d), C.byref(h)) == 1
    assert b"dt" in L.bm_last_error()
    assert L.bm_server_submit(None, None, None) == 1
    assert L.bm_server_destroy(None) == 0
