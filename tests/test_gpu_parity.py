"""HIP path (libbm_mpm.so, reference mode) against the oracle on identical inputs.

Bars (stated per test):
  * fixed-point node mass, active-node set and allocated-block set: bit-exact;
  * total grid mass every substep: exactly the sum of the particles' quanta;
  * P2G momentum / force: FP32 atomics accumulate in arbitrary order, so
    rtol 2e-5 of the field's largest magnitude;
  * G2P and couplers fed an identical grid: v, C and node velocities to 1 ulp
    level (rtol 1e-6), F within rtol 2e-5 (device logf/expf vs glibc);
  * multi-substep trajectories: rtol stated in each test.
"""
import numpy as np
import pytest

from tests._util import (assert_close_field, gpu_solver, ocollider, oracle_for, params_from, random_scene,
                         rel_err)

pytestmark = pytest.mark.gpu

DX = 1.0 / 32.0
ORG = (0, -4, 0)
DIMS = (24, 16, 24)


def _pair(seed, **kw):
    P = random_scene(seed, dx=DX, **kw)
    prm = params_from()
    g = gpu_solver(P, DX, ORG, DIMS, prm)
    o = oracle_for(P, DX, prm, g.mass_quantum_exp())
    return P, prm, g, o


def test_p2g_grid_matches_oracle():
    P, prm, g, o = _pair(1)
    g.p2g(1e-4)
    o.p2g()
    G = g.download_grid()
    O = o.grid_dense(g.origin, g.dims)
    assert np.array_equal(G["mass_q"], O["mass_q"]), "fixed-point node mass differs"
    assert np.array_equal(G["active"], O["active"]), "active node set differs"
    assert np.array_equal(G["alloc"], O["alloc"]), "allocated block set differs"
    assert_close_field(G["vel"], O["vel"], 2e-5, "P2G momentum")
    assert_close_field(G["force"], O["force"], 2e-5, "P2G stress force")
    assert g.grid_mass_total() == o.grid_mass_q_total()


def test_total_mass_is_particle_quanta():
    P, prm, g, o = _pair(2)
    e = g.mass_quantum_exp()
    q = np.rint(np.ldexp(P["m"].astype(np.float64), -e)).astype(np.int64)
    for _ in range(3):
        g.p2g(1e-4)
        assert g.grid_mass_total() == int(q.sum())
        g.grid_update(1e-4)
        g.g2p(1e-4)


def test_grid_update_matches_oracle():
    P, prm, g, o = _pair(3)
    g.p2g(1e-4)
    o.p2g()
    g.grid_update(1e-4)
    o.update(1e-4)
    G = g.download_grid()
    O = o.grid_dense(g.origin, g.dims)
    assert np.array_equal(G["active"], O["active"])
    assert_close_field(G["mass"], O["mass"], 0.0, "node mass")  # both convert the same integer
    assert_close_field(G["vel"], O["vel"], 5e-5, "updated node velocity")


def _identical_grid(seed):
    """GPU and oracle holding the same post-update grid (the oracle's, uploaded)."""
    P, prm, g, o = _pair(seed)
    o.p2g()
    o.update(1e-4)
    O = o.grid_dense(g.origin, g.dims)
    g.upload_grid(g.origin, O["mass"].astype(np.float32), O["vel"].astype(np.float32), O["active"])
    return P, prm, g, o, O


def test_g2p_from_identical_grid():
    P, prm, g, o, _ = _identical_grid(4)
    g.g2p(1e-4)
    o.g2p(1e-4)
    Pg = g.download()
    Po = o.get_particles()
    assert_close_field(Pg["v"], Po["v"], 1e-6, "particle v")
    assert_close_field(Pg["C"], Po["C"], 1e-6, "particle C")
    assert_close_field(Pg["x"], Po["x"], 1e-7, "particle x")
    assert_close_field(Pg["F"], Po["F"], 2e-5, "particle F")


@pytest.mark.parametrize("method,shape", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_couplers_from_identical_grid(method, shape):
    from basements_amd import Collider

    P, prm, g, o, O = _identical_grid(5)
    q = np.array([0.9, 0.1, 0.3, 0.2], np.float32)
    q = (q / np.linalg.norm(q)).astype(np.float32)
    col = Collider(center=(0.35, 0.12, 0.33), quat=tuple(float(v) for v in q),
                   half_extents=(0.06, 0.05, 0.06) if shape == 1 else (0.07, 0.05, 0.06),
                   lin_vel=(0.1, -0.5, 0.2), ang_vel=(1.0, 2.0, 3.0), friction=0.5, mass=2.0, shape=shape)
    fg, tg, ng = g.couple(method, col, 1e-4)
    fo, to, no = o.couple(method, ocollider(col), 1e-4)
    assert ng == no and ng > 0
    assert_close_field(fg, fo, 1e-4, "reaction force")
    assert_close_field(tg, to, 1e-4, "reaction torque")
    G = g.download_grid()
    O2 = o.grid_dense(g.origin, g.dims)
    assert np.array_equal(G["active"], O2["active"])
    assert_close_field(G["vel"], O2["vel"], 1e-6, "node velocity after coupling")
    assert_close_field(G["mass"], O2["mass"], 1e-7, "node mass after coupling")


def test_short_trajectory_matches_oracle():
    """20 coupled substeps (P2G, update, M1 wheel, G2P): rtol 1e-4 on x, v, F, C."""
    from basements_amd import Collider

    P, prm, g, o = _pair(6, strain=0.01, vel=0.05, cgrad=0.2)
    col = Collider(center=(0.35, 0.26, 0.35), quat=(0.70710677, 0.70710677, 0.0, 0.0), half_extents=(0.06, 0.03, 0.06),
                   lin_vel=(0.2, 0.0, 0.0), ang_vel=(0.0, 0.0, -2.0), friction=0.5, mass=5.0, shape=1)
    e = g.mass_quantum_exp()
    q = np.rint(np.ldexp(P["m"].astype(np.float64), -e)).astype(np.int64)
    dt = 1e-4
    for _ in range(20):
        g.p2g(dt)
        o.p2g()
        assert g.grid_mass_total() == o.grid_mass_q_total() == int(q.sum())
        g.grid_update(dt)
        o.update(dt)
        rg = g.couple(1, col, dt)
        ro = o.couple(1, ocollider(col), dt)
        assert rg[2] == ro[2]
        g.g2p(dt)
        o.g2p(dt)
    Pg = g.download()
    Po = o.get_particles()
    for k in ("x", "v", "F", "C"):
        assert rel_err(Pg[k], Po[k]) < 1e-4, (k, rel_err(Pg[k], Po[k]))


def test_out_of_domain_is_reported():
    from basements_amd import BMError

    P = random_scene(7, dx=DX)
    P["x"][0] = (5.0, 5.0, 5.0)
    g = gpu_solver(P, DX, ORG, DIMS, params_from())
    g.step(1e-4)
    with pytest.raises(BMError) as ei:
        g.synchronize()
    assert ei.value.code == 6
