"""Device body of the coupled loop: bm_mpm_step_coupled against the same loop
driven from the host (bm_mpm_couple + the oracle's TimeIntegrator::integrate),
the order of bench_compare_all.cpp:642-663 with P2G/update before and G2P
after the coupler (mpm_rigid_coupler.h:209-211)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from _util import gpu_solver, params_from, random_scene

pytestmark = pytest.mark.gpu

DX = 1.0 / 32.0
ORIGIN = (-4, -4, -4)
DIMS = (32, 20, 32)


def _f32(a):
    return np.asarray(a, np.float32)


def _bodies(method):
    from basements_amd import SHAPE_BOX, SHAPE_CYLINDER, Body

    if method == 0:
        # box pressed into the soil top, tilted, moving
        q = (math.cos(0.2), math.sin(0.2) * 0.6, math.sin(0.2) * 0.8, 0.0)
        return Body(position=(0.35, 0.22, 0.33), orientation=q, linear_velocity=(0.1, -0.4, 0.05),
                    angular_velocity=(0.2, 0.5, -0.3), inv_mass=1.0 / 2.0, linear_damping=0.0, angular_damping=0.0,
                    applied_force=(0.0, -9.81 * 2.0, 0.0), half_extents=(0.08, 0.05, 0.06), friction=0.5, mass=2.0,
                    shape=SHAPE_BOX)
    # rolling wheel: local Y (axle) -> world Z, as the rover-wheel scenario
    q = (math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0)
    return Body(position=(0.33, 0.25, 0.34), orientation=q, linear_velocity=(0.2, 0.0, 0.0),
                angular_velocity=(0.0, 0.0, -2.0), inv_mass=1.0 / 5.1, linear_damping=0.0, angular_damping=0.0,
                applied_force=(0.0, -50.0, 0.0), applied_torque=(0.0, 0.0, 0.3), half_extents=(0.1, 0.04, 0.1),
                friction=0.6, mass=5.1, shape=SHAPE_CYLINDER)


def _host_loop(S, body, method, dt, k):
    """The coupled loop with the body on the host: FP32 accumulation, the
    oracle's FP32 TimeIntegrator::integrate (pinned to the reference)."""
    from oracle.pyoracle import Body as OBody, OracleSim

    integ = OracleSim(DX, "f32")
    st = dict(x=_f32(body.position), q=_f32(body.orientation), v=_f32(body.linear_velocity),
              w=_f32(body.angular_velocity), f=_f32(body.force), t=_f32(body.torque))
    last = None
    for _ in range(k):
        S.p2g(dt)
        S.grid_update(dt)
        st["f"] = st["f"] + _f32(body.applied_force)
        st["t"] = st["t"] + _f32(body.applied_torque)
        from basements_amd import Collider

        col = Collider(center=st["x"], quat=st["q"], half_extents=body.half_extents, lin_vel=st["v"], ang_vel=st["w"],
                       friction=body.friction, mass=body.mass, shape=body.shape)
        f, t, n = S.couple(method, col, dt)
        last = (f, t, n)
        st["f"] = st["f"] + _f32(f)
        st["t"] = st["t"] + _f32(t)
        ob = OBody(x=tuple(st["x"]), quat=tuple(st["q"]), v=tuple(st["v"]), w=tuple(st["w"]),
                   inv_mass=float(np.float32(body.inv_mass)), inv_I=np.asarray(body.inv_inertia).reshape(3, 3),
                   lin_damp=body.linear_damping, ang_damp=body.angular_damping, force=tuple(st["f"]),
                   torque=tuple(st["t"]))
        ob = integ.integrate_body(ob, dt)
        st = dict(x=_f32(ob.x), q=_f32(ob.quat), v=_f32(ob.v), w=_f32(ob.w), f=_f32(ob.force), t=_f32(ob.torque))
        S.g2p(dt)
    return st, last


def _grid_state(seed):
    """A deterministic updated grid: a slab of massive nodes with random
    velocities under the body, empty nodes around (M1 claims those)."""
    rng = np.random.default_rng(seed)
    dims = np.array(DIMS)
    mass = np.zeros(dims, np.float32)
    vel = np.zeros(tuple(dims) + (3,), np.float32)
    act = np.zeros(dims, np.uint8)
    lo = np.array([6, 4, 6]) - np.array(ORIGIN)
    hi = np.array([20, 12, 20]) - np.array(ORIGIN)
    sl = tuple(slice(a, b) for a, b in zip(lo, hi))
    mass[sl] = rng.uniform(1e-4, 2e-3, size=tuple(hi - lo)).astype(np.float32)
    vel[sl] = rng.normal(0.0, 0.3, size=tuple(hi - lo) + (3,)).astype(np.float32)
    act[sl] = 1
    return mass, vel, act


def _grid_solver(mode, seed):
    from basements_amd import Solver

    S = Solver(DX, ORIGIN, DIMS, 16, mode)
    mass, vel, act = _grid_state(seed)
    S.upload_grid(ORIGIN, mass, vel, act)
    return S


@pytest.mark.parametrize("method", [0, 1], ids=["M0", "M1"])
def test_couple_body_equals_couple_on_the_same_grid(method):
    """Same grid bits: the device-built collider and the body's reduction give
    the reaction and node velocities of bm_mpm_couple exactly, and the body
    ends where the oracle's FP32 integrator puts it."""
    from oracle.pyoracle import Body as OBody, OracleSim

    dt = 2e-4
    body = _bodies(method)
    body.position = (0.42, 0.30, 0.41) if method == 0 else (0.42, 0.34, 0.41)
    A, B = _grid_solver(0, 3 + method), _grid_solver(0, 3 + method)
    f, t, n = A.couple(method, body.collider(), dt)
    B.set_body(body)
    gf, gt, gn = B.couple_body(method, dt, wait=True)
    assert n > 0, "the body touched no node: the test would pin nothing"
    np.testing.assert_array_equal(_f32(gf), _f32(f))
    np.testing.assert_array_equal(_f32(gt), _f32(t))
    assert gn == n
    ga, gb = A.download_grid(), B.download_grid()
    for k in ("vel", "mass", "active"):
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    ob = OBody(x=tuple(_f32(body.position)), quat=tuple(_f32(body.orientation)), v=tuple(_f32(body.linear_velocity)),
               w=tuple(_f32(body.angular_velocity)), inv_mass=float(np.float32(body.inv_mass)),
               inv_I=np.asarray(body.inv_inertia).reshape(3, 3), lin_damp=body.linear_damping,
               ang_damp=body.angular_damping, force=tuple(_f32(body.applied_force) + _f32(f)),
               torque=tuple(_f32(body.applied_torque) + _f32(t)))
    ob = OracleSim(DX, "f32").integrate_body(ob, dt)
    got, (rf, rt, rn) = B.get_body()
    np.testing.assert_array_equal(_f32(got.position), _f32(ob.x))
    np.testing.assert_array_equal(_f32(got.orientation), _f32(ob.quat))
    np.testing.assert_array_equal(_f32(got.linear_velocity), _f32(ob.v))
    np.testing.assert_array_equal(_f32(got.angular_velocity), _f32(ob.w))
    np.testing.assert_array_equal(_f32(got.force), 0.0)
    np.testing.assert_array_equal(_f32(rf), _f32(f))
    assert rn == n


@pytest.mark.parametrize("mode", [0, 1], ids=["reference", "fast"])
@pytest.mark.parametrize("method", [0, 1], ids=["M0", "M1"])
def test_step_coupled_tracks_host_loop(mode, method):
    """Twelve coupled substeps on the device against the loop driven from the
    host.  Node momenta are FP32 sums whose order follows the atomics, so two
    runs of either loop agree to rounding, not bit for bit: the bar is a
    relative 1e-4 of each field's scale."""
    from _util import assert_close_field

    P = random_scene(21 + method, dx=DX, cells_lo=(6, 0, 6), cells_hi=(16, 8, 16), vel=0.2, strain=0.02, cgrad=1.0)
    prm = params_from()
    dt, k = 2e-4, 12
    body = _bodies(method)
    A = gpu_solver(P, DX, ORIGIN, DIMS, prm, mode=mode)
    B = gpu_solver(P, DX, ORIGIN, DIMS, prm, mode=mode)
    want, (f, t, n) = _host_loop(A, body, method, dt, k)
    B.set_body(body)
    B.step_coupled(method, dt, k)
    got, (gf, gt, gn) = B.get_body()
    assert n > 0, "the body never touched the soil: the test would pin nothing"
    assert_close_field(got.position, want["x"], 1e-6, "position")
    assert_close_field(got.orientation, want["q"], 1e-6, "orientation")
    assert_close_field(got.linear_velocity, want["v"], 1e-4, "velocity")
    assert_close_field(got.angular_velocity, want["w"], 1e-4, "angular velocity")
    assert_close_field(gf, f, 2e-3, "reaction force")
    pa, pb = A.download(), B.download()
    for key in ("x", "v", "F", "C"):
        assert_close_field(pb[key], pa[key], 1e-4, key)


def test_couple_body_waits_and_static_body():
    """couple_body with a host reaction equals get_body's; a body with
    integrate = 0 keeps its pose and accumulates its forces."""
    P = random_scene(5, dx=DX, cells_lo=(6, 0, 6), cells_hi=(16, 8, 16))
    prm = params_from()
    body = _bodies(0)
    body.integrate = False
    S = gpu_solver(P, DX, ORIGIN, DIMS, prm, mode=1)
    S.set_body(body)
    S.p2g(1e-4)
    S.grid_update(1e-4)
    f, t, n = S.couple_body(0, 1e-4, wait=True)
    got, (gf, gt, gn) = S.get_body()
    assert n > 0 and gn == n
    np.testing.assert_array_equal(gf, f)
    np.testing.assert_array_equal(_f32(got.position), _f32(body.position))
    np.testing.assert_array_equal(_f32(got.force), _f32(body.applied_force) + _f32(f))


def test_couple_body_state_errors():
    from basements_amd import BMError

    P = random_scene(6, dx=DX, cells_lo=(6, 0, 6), cells_hi=(10, 4, 10))
    S = gpu_solver(P, DX, ORIGIN, DIMS, params_from(), mode=1)
    S.p2g(1e-4)
    S.grid_update(1e-4)
    with pytest.raises(BMError, match="STATE"):
        S.couple_body(0, 1e-4)
    body = _bodies(0)
    S.set_body(body)
    with pytest.raises(BMError, match="INVALID_ARG"):
        S.couple_body(7, 1e-4)
    body.shape = 9
    with pytest.raises(BMError, match="INVALID_ARG"):
        S.set_body(body)
