"""The reference-plugin form on the device: an engine's view of
libbm_mpm_soil_plugin.so (basements_amd/plugin/bm_mpm_soil_plugin.c, built
against include/basements/engine/plugins/plugin_capi.h where the reference's
headers are present).  on_load creates the soil handle; the engine fills it
through the plugin's handle and gives it a tool body; eight engine frames of
on_pre_step / on_post_step must leave soil and body where eight coupled
substeps of a handle of the same configuration, driven directly, leave them;
on_unload releases the handle.  Skipped when the plugin library was not
built."""
from __future__ import annotations

import os

import numpy as np
import pytest

from basements_amd import build as B

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(B.PLUGIN_LIB), reason="plugin library not built")]

DX = 1.0 / 120.0  # the plugin's compiled-in domain: 256^3 nodes from node 0
NODES = (256, 256, 256)
CAP = 1 << 23
DT = 1e-4


def _scene():
    from basements_amd import BODY_DYNAMIC, SHAPE_BOX, Body
    from basements_amd.scenes import soil_block

    P = soil_block(DX, (100, 2, 100), (156, 18, 156), seed=11, rho=1600.0)
    body = Body(position=(128 * DX, 21.0 * DX, 128 * DX), linear_velocity=(0.0, -0.5, 0.0), inv_mass=0.5,
                half_extents=(5 * DX, 3 * DX, 5 * DX), mass=2.0, integrate=BODY_DYNAMIC, shape=SHAPE_BOX)
    return P, body


def test_plugin_frames_equal_direct_coupled_substeps():
    import ctypes as C

    from basements_amd import M1, MODE_FAST, Solver
    from _plugin import Plugin

    P, body = _scene()
    p = Plugin(B.PLUGIN_LIB)
    assert p.on_load(None) == 0, p.status()
    try:
        h = p.handle()
        assert h
        S = Solver.attach(h, DX, (0, 0, 0), NODES, CAP, mode=MODE_FAST)
        S.upload(P)
        bc = body.to_c()
        assert p.set_body(C.addressof(bc)) == 0
        for _ in range(8):
            p.on_pre_step(DT)
            p.on_post_step(DT)
        assert p.status() == 0, p.status()
        got = S.download()
        gbody, (_, _, n_in) = S.get_body()
        S.close()  # not the owner: the handle stays the plugin's
        assert p.handle() == h
    finally:
        p.on_unload()
    assert p.handle() is None

    D = Solver(DX, (0, 0, 0), NODES, CAP, mode=MODE_FAST)
    D.upload(P)
    D.set_body(body)
    D.step_coupled(M1, DT, 8)
    want = D.download()
    wbody, (_, _, w_in) = D.get_body()
    D.close()

    assert n_in == w_in and n_in > 0, (n_in, w_in)
    for k in ("x", "v", "F", "C"):
        scale = float(np.abs(want[k]).max())
        err = float(np.abs(got[k] - want[k]).max())
        assert err <= 1e-5 * scale + 1e-7, (k, err, scale)
    assert np.array_equal(got["m"], want["m"]) and np.array_equal(got["V0"], want["V0"])
    assert np.abs(np.asarray(gbody.position) - np.asarray(wbody.position)).max() <= 1e-6
    assert float(np.abs(got["x"] - P["x"]).max()) > 0
