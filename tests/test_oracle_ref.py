"""Pin the oracle's FP32 restatement bit-for-bit to the reference itself.

oracle/_ref/libref_mpm.so is the reference's own MPMSolver / SVD / M0 / M1 /
TimeIntegrator compiled unchanged from its sources by oracle/Makefile; it is
only present where the reference tree was available at build time, so these
tests skip elsewhere (tests/test_golden.py carries the same pins as committed
fixtures).
"""
import numpy as np
import pytest

from oracle.pyoracle import (Body, Collider, OracleSim, Params, RefSim, make_particles, ref_available, round32)

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")


def _mats(rng, n):
    out = [np.eye(3), np.diag([1.5, 0.5, 1.5]), np.diag([1.0, 0.9, 1.0]), 0.95 * np.eye(3), np.zeros((3, 3)),
           np.diag([1e-7, 1.0, 2.0]), np.array([[1, 1e-7, 0], [0, 1, 0], [0, 0, 1.0]])]
    for k in range(n):
        if k % 3 == 0:
            out.append(np.eye(3) + 1e-3 * rng.normal(size=(3, 3)))
        elif k % 3 == 1:
            out.append(rng.normal(size=(3, 3)))
        else:
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            out.append(q @ np.diag(rng.uniform(0.5, 1.5, 3)) @ q.T)
    return [np.asarray(m, np.float32).astype(np.float64) for m in out]


def test_svd_bit_exact():
    rng = np.random.default_rng(0)
    o = OracleSim(0.1)
    r = RefSim(0.1)
    for F in _mats(rng, 600):
        a, b = o.svd(F), r.svd(F)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), F


def _scene(seed, n=400, dx=1 / 64):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.1, 0.5, size=(n, 3))
    P = round32(make_particles(x, 0.1 * rng.normal(size=(n, 3)), mass=7.6e-4, rho=1600))
    P["F"] = np.asarray(np.eye(3) + 0.03 * rng.normal(size=(n, 3, 3)), np.float32).astype(np.float64)
    P["C"] = np.asarray(0.5 * rng.normal(size=(n, 3, 3)), np.float32).astype(np.float64)
    prm = Params(E=3.5e5, nu=0.3, rho=1600, alpha=0.386, mu_floor=0.4)
    o, r = OracleSim(dx, params=prm), RefSim(dx, params=prm)
    o.set_particles(P)
    r.set_particles(P)
    return o, r


ORG, DIMS = (-2, -2, -2), (40, 40, 40)


def _grid_equal(o, r):
    go, gr = o.grid_dense(ORG, DIMS), r.grid_dense(ORG, DIMS)
    for k in ("mass", "vel", "force", "active", "alloc"):
        assert np.array_equal(go[k], gr[k]), k
    assert o.num_blocks() == r.num_blocks()


def test_phases_bit_exact():
    o, r = _scene(1)
    o.p2g()
    r.p2g()
    _grid_equal(o, r)
    o.update(1e-4)
    r.update(1e-4)
    _grid_equal(o, r)
    o.g2p(1e-4)
    r.g2p(1e-4)
    po, pr = o.get_particles(), r.get_particles()
    for k in "xvFC":
        assert np.array_equal(po[k], pr[k]), k


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("shape", [0, 1])
def test_couplers_bit_exact(method, shape):
    o, r = _scene(2)
    o.p2g()
    r.p2g()
    o.update(1e-4)
    r.update(1e-4)
    q = np.array([0.9, 0.1, 0.3, 0.2])
    q = np.asarray(q / np.linalg.norm(q), np.float32).astype(np.float64)
    col = Collider(center=(0.3, 0.3, 0.3), quat=tuple(q), half_extents=(0.05, 0.08, 0.06), lin_vel=(0.1, -0.5, 0.2),
                   ang_vel=(1, 2, 3), friction=0.5, mass=2.0, shape=shape)
    ro, rr = o.couple(method, col, 1e-4), r.couple(method, col, 1e-4)
    assert ro[2] == rr[2] and ro[2] > 0
    assert np.array_equal(ro[0], rr[0]) and np.array_equal(ro[1], rr[1])
    _grid_equal(o, r)


def test_trajectory_bit_exact():
    o, r = _scene(3, n=300)
    o.step(1e-4, 100)
    r.step(1e-4, 100)
    po, pr = o.get_particles(), r.get_particles()
    for k in "xvFC":
        assert np.array_equal(po[k], pr[k]), k


def test_rigid_integrator_bit_exact():
    o = OracleSim(0.1)
    r = RefSim(0.1)
    b = Body(x=(0.1, 0.5, -0.2), quat=(0.9, 0.1, 0.3, 0.2), v=(0.3, -1.0, 0.2), w=(1.0, -2.0, 0.5), inv_mass=0.5,
             inv_I=np.diag([2.0, 3.0, 4.0]), lin_damp=0.01, ang_damp=0.02, force=(1.0, 2.0, -3.0),
             torque=(0.1, 0.2, 0.3))
    for _ in range(20):
        bo, br = o.integrate_body(b, 1 / 240), r.integrate_body(b, 1 / 240)
        assert np.array_equal(bo.pack(), br.pack())
        b = Body(x=bo.x, quat=bo.quat, v=bo.v, w=bo.w, inv_mass=b.inv_mass, inv_I=b.inv_I, lin_damp=b.lin_damp,
                 ang_damp=b.ang_damp, force=(1.0, 2.0, -3.0), torque=(0.1, 0.2, 0.3))


@pytest.mark.parametrize("method", [0, 1])
def test_coupled_loop_bit_exact(method):
    """The oracle's coupled loop (orc_*_step_coupled) against the reference's
    own phases, coupler and integrator composed in the contract order of
    mpm_rigid_coupler.h:209-211 / bench_compare_all.cpp:642-663, with the
    body's force accumulated in FP32 as RigidBody holds it."""
    o, r = _scene(4, n=500)
    f32 = lambda a: np.asarray(a, np.float32)  # noqa: E731
    col = Collider(half_extents=(0.05, 0.05, 0.05), friction=0.5, mass=1.0, shape=0)
    body = Body(x=(0.3, 0.36, 0.3), v=(0.05, -1.4, 0.0), w=(0.3, 0.0, -0.2), inv_mass=1.0, lin_damp=0.0,
                ang_damp=0.0)
    got, rec = o.step_coupled(method, col, body, (0.0, -9.81, 0.0), (0.0, 0.0, 0.0), 1e-4, 30)
    b = body
    touched = 0
    for k in range(30):
        f = f32(b.force) + f32((0.0, -9.81, 0.0))
        t = f32(b.torque) + f32((0.0, 0.0, 0.0))
        r.p2g()
        r.update(1e-4)
        c = Collider(center=b.x, quat=b.quat, half_extents=col.half_extents, lin_vel=b.v, ang_vel=b.w,
                     friction=col.friction, mass=col.mass, shape=col.shape)
        rf, rt, n = r.couple(method, c, 1e-4)
        touched += n
        b = Body(x=b.x, quat=b.quat, v=b.v, w=b.w, inv_mass=b.inv_mass, inv_I=b.inv_I, lin_damp=0.0, ang_damp=0.0,
                 force=tuple(f + f32(rf)), torque=tuple(t + f32(rt)))
        b = r.integrate_body(b, 1e-4)
        r.g2p(1e-4)
        want = np.concatenate([b.x, b.quat, b.v, b.w, rf, rt, [n]])
        assert np.array_equal(rec[k], want), k
    assert touched > 0, "the body never touched the soil"
    po, pr = o.get_particles(), r.get_particles()
    for k in "xvFC":
        assert np.array_equal(po[k], pr[k]), k


def test_coupled_loop_kinematic_is_dynamic_without_inverse_mass():
    """A kinematic body moves at its own velocities: bit-identical to the
    dynamic integration with zero inverse mass and inertia and no damping."""
    runs = []
    for kin in (True, False):
        o, _ = _scene(5, n=300)
        col = Collider(half_extents=(0.06, 0.04, 0.05), friction=0.4, mass=2.0, shape=1)
        body = Body(x=(0.3, 0.35, 0.3), quat=(0.9, 0.1, 0.3, 0.2), v=(0.3, -0.2, 0.1), w=(0.0, 0.0, -3.0),
                    inv_mass=1.0 if kin else 0.0, inv_I=np.eye(3) if kin else np.zeros((3, 3)), lin_damp=0.0,
                    ang_damp=0.0)
        runs.append(o.step_coupled(1, col, body, (0.0, -20.0, 0.0), (0.0, 0.0, 0.0), 1e-4, 20, kinematic=kin)[1])
    assert np.array_equal(runs[0], runs[1])
    assert runs[0][:, 19].max() > 0
