"""Batches of scenes (bm_batch_*): every scene of a batch against the same
scene run alone on a single-scene fast handle, through scene removal, graph
replay and autotuning.

Node momenta are FP32 sums whose order follows the atomics, so a scene in a
batch and the same scene alone agree to rounding, not bit for bit; the bar is
a relative 1e-4 of each field's scale after the stated substeps (the same bar
as the coupled-loop tests).  Grid mass is fixed point and order-free."""
from __future__ import annotations

import math

import numpy as np
import pytest

from _util import assert_close_field, gpu_solver, params_from, random_scene

pytestmark = pytest.mark.gpu

DT = 2e-4


def _scenes():
    """Three uneven scenes: a box on M0, a rolling wheel on M1, no body."""
    from basements_amd import BODY_DYNAMIC, BODY_KINEMATIC, M0, M1, SHAPE_BOX, SHAPE_CYLINDER, Body, SceneDesc

    out = []
    # 1: dx 1/32, box pressed in, M0
    P = random_scene(31, dx=1 / 32, cells_lo=(6, 0, 6), cells_hi=(16, 8, 16), vel=0.2, strain=0.02, cgrad=1.0)
    prm = params_from(E=3.5e5, alpha=0.386)
    q = (math.cos(0.2), math.sin(0.2) * 0.6, math.sin(0.2) * 0.8, 0.0)
    body = Body(position=(0.35, 0.22, 0.33), orientation=q, linear_velocity=(0.1, -0.4, 0.05),
                angular_velocity=(0.2, 0.5, -0.3), inv_mass=0.5, linear_damping=0.0, angular_damping=0.0,
                applied_force=(0.0, -9.81 * 2.0, 0.0), half_extents=(0.08, 0.05, 0.06), friction=0.5, mass=2.0,
                shape=SHAPE_BOX, integrate=BODY_DYNAMIC)
    out.append((P, prm, SceneDesc(dx=1 / 32, origin=(-4, -4, -4), dims=(32, 20, 32), E=prm.E, nu=prm.nu, rho=prm.rho,
                                  alpha=prm.alpha, mu_floor=prm.mu_floor, coupler=M0), body))
    # 2: dx 1/40, a dragged wheel on M1 (kinematic)
    P = random_scene(32, dx=1 / 40, cells_lo=(8, 0, 8), cells_hi=(20, 8, 18), vel=0.1, strain=0.01, cgrad=0.5)
    prm = params_from(E=8e5, alpha=0.30, mu_floor=0.5)
    body = Body(position=(0.3, 0.25, 0.32), orientation=(math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0),
                linear_velocity=(0.3, 0.0, 0.0), angular_velocity=(0.0, 0.0, -3.0), inv_mass=1 / 5.0,
                linear_damping=0.0, angular_damping=0.0, applied_force=(0.0, -49.0, 0.0),
                half_extents=(0.07, 0.03, 0.07), friction=0.6, mass=5.0, shape=SHAPE_CYLINDER,
                integrate=BODY_KINEMATIC)
    out.append((P, prm, SceneDesc(dx=1 / 40, origin=(-4, -4, -4), dims=(32, 24, 32), E=prm.E, nu=prm.nu, rho=prm.rho,
                                  alpha=prm.alpha, mu_floor=prm.mu_floor, coupler=M1), body))
    # 3: dx 1/24, free soil column, no body
    P = random_scene(33, dx=1 / 24, cells_lo=(4, 0, 4), cells_hi=(10, 10, 9), vel=0.3, strain=0.03, cgrad=1.5)
    prm = params_from(E=2e5, alpha=0.42)
    out.append((P, prm, SceneDesc(dx=1 / 24, origin=(-4, -4, -4), dims=(20, 24, 20), E=prm.E, nu=prm.nu, rho=prm.rho,
                                  alpha=prm.alpha, mu_floor=prm.mu_floor), None))
    return out


def _alone(P, prm, sd, body, k):
    """The scene on its own fast handle: step_coupled with a body, step without."""
    S = gpu_solver(P, sd.dx, sd.origin, sd.dims, prm, mode=1)
    if body is not None:
        S.set_body(body)
        S.step_coupled(sd.coupler, DT, k)
        got = S.get_body()[0]
    else:
        S.step(DT, k)
        got = None
    return S.download(), got


def _batch(scenes, graph=False):
    from basements_amd import Batch

    B = Batch(max_particles=200_000, max_blocks=40_000, max_scenes=8)
    B.set_graph(graph)
    ids = [B.add_scene(sd, P, body) for (P, prm, sd, body) in scenes]
    return B, ids


def _same(got, want, what, rtol=1e-4):
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], want[key], rtol, f"{what} {key}")


def _same_body(got, want, what):
    assert_close_field(got.position, want.position, 1e-5, f"{what} body position")
    assert_close_field(got.linear_velocity, want.linear_velocity, 1e-3, f"{what} body velocity")


def test_each_scene_in_a_batch_matches_it_alone():
    scenes = _scenes()
    B, ids = _batch(scenes)
    assert B.usage()[0] == 3
    k = 20
    B.step(DT, k)
    for sid, (P, prm, sd, body) in zip(ids, scenes):
        want, wbody = _alone(P, prm, sd, body, k)
        _same(B.download_scene(sid), want, f"scene {sid}")
        if body is not None:
            _same_body(B.get_body(sid)[0], wbody, f"scene {sid}")
    # the kinematic wheel moved at its own velocity
    b2 = B.get_body(ids[1])[0]
    assert b2.position[0] == pytest.approx(0.3 + 0.3 * DT * k, abs=1e-6)


def test_removing_a_scene_keeps_the_others():
    scenes = _scenes()
    B, ids = _batch(scenes)
    B.step(DT, 8)
    B.remove_scene(ids[1])
    assert B.usage()[0] == 2
    B.step(DT, 8)
    for sid, j in ((ids[0], 0), (ids[2], 2)):
        P, prm, sd, body = scenes[j]
        want, wbody = _alone(P, prm, sd, body, 16)
        _same(B.download_scene(sid), want, f"scene {sid} after removal")
    # a new scene takes the freed room and runs like it alone
    P, prm, sd, body = scenes[1]
    sid = B.add_scene(sd, P, body)
    B.step(DT, 10)
    want, wbody = _alone(P, prm, sd, body, 10)
    _same(B.download_scene(sid), want, "readmitted scene")
    _same_body(B.get_body(sid)[0], wbody, "readmitted scene")


def test_graph_replays_match_eager_steps():
    """Replays of the captured substeps, with eager work (downloads) between
    them, against the same batch stepped eagerly."""
    scenes = _scenes()
    E, eids = _batch(scenes, graph=False)
    G, gids = _batch(scenes, graph=True)
    E.set_tuning(0, 4, 0)  # short re-binning cycle: 2 x 4 = 8 substeps per replay
    G.set_tuning(0, 4, 0)
    for rnd in range(3):
        E.step(DT, 12)
        G.step(DT, 12)
        for a, b in zip(eids, gids):
            _same(G.download_scene(b), E.download_scene(a), f"round {rnd} scene {b}")


def test_admission_by_capacity():
    from basements_amd import Batch, BMError

    scenes = _scenes()
    B = Batch(max_particles=scenes[0][0]["x"].shape[0] + 10, max_blocks=40_000, max_scenes=4)
    B.add_scene(scenes[0][2], scenes[0][0], scenes[0][3])
    with pytest.raises(BMError, match="OOM"):
        B.add_scene(scenes[2][2], scenes[2][0])
    B2 = Batch(max_particles=200_000, max_blocks=10, max_scenes=4)
    with pytest.raises(BMError, match="OOM"):
        B2.add_scene(scenes[2][2], scenes[2][0])


def test_autotune_keeps_the_scenes_state():
    scenes = _scenes()
    A, aids = _batch(scenes)
    T, tids = _batch(scenes)
    unit, sort_every, groups = T.autotune(DT, substeps=8)
    assert unit in (512, 640) and sort_every in (8, 16, 32) and groups in (2048, 4096, 8192)
    A.step(DT, 10)
    T.step(DT, 10)
    for a, b in zip(aids, tids):
        _same(T.download_scene(b), A.download_scene(a), f"tuned scene {b}")
