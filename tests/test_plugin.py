"""The reference-plugin form of the substep (basements_amd/plugin/
bm_mpm_soil_plugin.c) against the reference's own plugin header
(include/basements/engine/plugins/plugin_capi.h): it compiles warning-free as
C99 and as C++17, exports basements_plugin_register, and registers the
callbacks an engine binds (plugin_capi.h:52-62, 92-102) with a valid info
record; before on_load its own exports refuse work and the step callbacks do
nothing.  Host CPU only (no callback here touches a GPU); skipped where the
reference's headers are not present."""
from __future__ import annotations

import os
import subprocess

import pytest

from basements_amd import build as B

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(B.REF_INCLUDE, B.PLUGIN_HDR)),
                                reason="the reference's plugin header is not present")


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_plugin_compiles_against_reference_header(tmp_path, lang):
    if not os.path.exists(B.LIB):
        pytest.skip("libbm_mpm.so not built")
    out = str(tmp_path / f"plugin_{lang.replace('+', 'x')}.so")
    r = subprocess.run(B.plugin_cmd(out, lang=lang), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    syms = subprocess.run(["nm", "-D", "--defined-only", out], capture_output=True, text=True).stdout
    for name in ("basements_plugin_register", "bm_soil_plugin_handle", "bm_soil_plugin_status",
                 "bm_soil_plugin_set_body"):
        assert f" T {name}" in syms, (name, syms)


def test_plugin_registers_and_is_inert_before_load(tmp_path):
    if not os.path.exists(B.LIB):
        pytest.skip("libbm_mpm.so not built")
    out = str(tmp_path / "plugin.so")
    r = subprocess.run(B.plugin_cmd(out), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    from _plugin import Plugin

    p = Plugin(out)
    assert p.cb.get_info and p.cb.on_load and p.cb.on_unload and p.cb.on_pre_step and p.cb.on_post_step
    assert not p.cb.on_render
    info = p.get_info()
    assert info.name == b"bm_mpm_soil" and info.version == b"1" and info.api_version == 1
    assert p.handle() is None and p.status() == 0
    assert p.set_body(None) == 7  # BM_ERR_STATE: no handle before on_load
    p.on_pre_step(1e-4)
    p.on_post_step(1e-4)
    p.on_unload()
    assert p.handle() is None and p.status() == 0
