"""The oracle against the committed golden fixtures (tests/golden/), written
from the reference's own build by tests/golden/make_golden.py, and against the
reference's committed repose-pile run.  These hold wherever the reference tree
is absent; tests/test_oracle_ref.py checks the same against the live build."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402  (inputs are rebuilt from the generator's seeds)
from oracle.pyoracle import Body, OracleSim, Params  # noqa: E402

GOLD = os.path.join(HERE, "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _sim():
    return OracleSim(G.SCENE["dx"], "f32", params=Params(**G.PARAMS))


def _grid(o):
    g = o.grid_dense(G.BOX_ORG, G.BOX_DIMS)
    return {"mass": g["mass"], "vel": g["vel"], "force": g["force"], "active": g["active"].astype(np.uint8),
            "alloc": g["alloc"].astype(np.uint8)}


def _eq(got, want, what):
    got = np.asarray(got, np.float64)
    want = np.asarray(want).astype(np.float64)
    assert np.array_equal(got, want), f"{what}: {int((got != want).sum())} entries differ"


def test_svd_matches_reference_fixture():
    d = _load("ref_svd.npz")
    o = _sim()
    for F, U, S, V in zip(d["F"], d["U"], d["S"], d["V"]):
        u, s, v = o.svd(F.astype(np.float64))
        _eq(u, U, "U")
        _eq(s, S, "sigma")
        _eq(v, V, "V")


def test_phases_match_reference_fixture():
    d = _load("ref_phases.npz")
    o = _sim()
    o.set_particles(G.scene(1))
    o.p2g()
    for k, v in _grid(o).items():
        _eq(v, d[f"p2g_{k}"], f"p2g {k}")
    o.update(G.SCENE["dt"])
    for k, v in _grid(o).items():
        _eq(v, d[f"update_{k}"], f"update {k}")
    o.g2p(G.SCENE["dt"])
    P = o.get_particles()
    for k in "xvFC":
        _eq(P[k], d[f"g2p_{k}"], f"g2p {k}")


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("shape", [0, 1])
def test_couplers_match_reference_fixture(method, shape):
    d = _load("ref_couplers.npz")
    o = _sim()
    o.set_particles(G.scene(2))
    o.p2g()
    o.update(G.SCENE["dt"])
    f, t, n = o.couple(method, G.colliders()[shape], G.SCENE["dt"])
    tag = f"m{method}_s{shape}"
    assert n == int(d[f"{tag}_nodes"][0]) and n > 0
    _eq(np.concatenate([f, t]), d[f"{tag}_reaction"], "reaction")
    g = _grid(o)
    _eq(g["vel"], d[f"{tag}_vel"], "node velocity")
    _eq(g["mass"], d[f"{tag}_mass"], "node mass")


def test_trajectory_matches_reference_fixture():
    d = _load("ref_trajectory.npz")
    o = _sim()
    o.set_particles(G.scene(3))
    o.step(G.SCENE["dt"], 50)
    P = o.get_particles()
    for k in "xvFC":
        _eq(P[k], d[k], k)


def test_integrator_matches_reference_fixture():
    d = _load("ref_body.npz")["seq"]
    o = _sim()
    b = G.body0()
    for want in d:
        b2 = o.integrate_body(b, 1.0 / 240.0)
        _eq(b2.pack(), want, "body")
        b = Body(x=b2.x, quat=b2.quat, v=b2.v, w=b2.w, inv_mass=b.inv_mass, inv_I=b.inv_I, lin_damp=b.lin_damp,
                 ang_damp=b.ang_damp, force=(1.0, 2.0, -3.0), torque=(0.1, 0.2, 0.3))


def test_repose_pile_reproduces_the_reference_run():
    """The oracle's FP32 MPMSolver::step on the reference's repose-pile column
    against the numbers the reference committed: the time series every 24
    substeps and the final summary, to the printed precision (the committed
    run came from another compiler, so bits are not expected)."""
    from _pile import column, measure, reference_run
    from oracle.pyoracle import make_particles

    ref = reference_run()
    sc = ref["scenario"]
    prm = Params(E=sc["E"], nu=sc["nu"], rho=sc["rho"], alpha=sc["alpha"], mu_floor=sc["mu_floor"],
                 gravity=(0.0, -sc["gravity"], 0.0))
    o = OracleSim(sc["dx"], "f32", params=prm)
    x = column(sc)
    o.set_particles(make_particles(x.astype(np.float64), mass=float(np.float32(sc["mass"])), rho=sc["rho"]))
    dt = float(np.float32(sc["dt"]))
    rows = {round(r[0], 4): r for r in ref["series"]}
    checked = 0
    for i in range(sc["steps"]):
        o.step(dt, 1)
        if i % 24 == 0:
            t = round(float(np.float32(i + 1) * np.float32(dt)), 4)
            P = o.get_particles()
            m = measure(P["x"], P["v"], sc)
            r = rows[t]
            assert m["height"] == pytest.approx(r[1], rel=2e-4, abs=1e-6), (t, "height")
            assert m["radius"] == pytest.approx(r[2], rel=2e-4, abs=1e-6), (t, "radius")
            assert m["angle_slope"] == pytest.approx(r[3], abs=0.05), (t, "slope angle")
            assert abs(m["at_rest"] - r[4]) <= 2, (t, "at rest")
            checked += 1
    assert checked == len(ref["series"])
    P = o.get_particles()
    m = measure(P["x"], P["v"], sc)
    s = ref["summary"]
    assert m["height"] == pytest.approx(s["settled_height_m"], rel=2e-4)
    assert m["radius"] == pytest.approx(s["settled_radius_m"], rel=2e-4)
    assert m["angle_geom"] == pytest.approx(s["angle_geom_deg"], abs=0.01)
    assert m["angle_slope"] == pytest.approx(s["angle_slope_deg"], abs=0.05)
    assert abs(m["at_rest"] - s["particles_at_rest"]) <= 2
