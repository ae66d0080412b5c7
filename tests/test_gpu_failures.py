"""Failure detection at the boundary: non-finite state is refused at upload
(BM_ERR_NONFINITE) and leaves the handle usable; a non-finite value that
reaches a particle on the device is flagged by the substep, not propagated
silently; a scene of a batch that goes non-finite fails alone."""
from __future__ import annotations

import numpy as np
import pytest

from _util import gpu_solver, params_from, random_scene

pytestmark = pytest.mark.gpu

DX = 1.0 / 32.0
ORG = (0, -4, 0)
DIMS = (24, 16, 24)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("field", ["x", "v", "F", "m"])
def test_non_finite_upload_is_refused(mode, field):
    from basements_amd import BMError

    P = random_scene(41, dx=DX)
    bad = {k: np.array(v, copy=True) for k, v in P.items()}
    bad[field].reshape(bad[field].shape[0], -1)[3, 0] = np.inf if field == "m" else np.nan
    S = gpu_solver(P, DX, ORG, DIMS, params_from(), mode=mode)
    with pytest.raises(BMError) as ei:
        S.upload(bad)
    assert ei.value.code == 4
    S.upload(P)  # still usable
    S.step(1e-4, 2)
    S.synchronize()


def test_non_finite_node_velocity_is_flagged():
    """A NaN seeded into a node that particles gather from: the G2P raises
    BM_ERR_NONFINITE at the next synchronisation."""
    from basements_amd import BMError

    P = random_scene(42, dx=DX)
    S = gpu_solver(P, DX, ORG, DIMS, params_from(), mode=0)
    org = np.array(ORG, np.int32)
    mass = np.full(DIMS, 1e-3, np.float32)  # node (i, j, k) of the box at [i, j, k]
    vel = np.zeros(DIMS + (3,), np.float32)
    active = np.ones(DIMS, np.uint8)
    # the node under particle 0's stencil centre
    i, j, k = (np.floor(P["x"][0] / DX - 0.5).astype(int) + 1) - org
    vel[i, j, k] = (np.nan, 0.0, 0.0)
    S.upload_grid(org, mass, vel, active)
    S.g2p(1e-4)
    with pytest.raises(BMError) as ei:
        S.synchronize()
    assert ei.value.code == 4


def test_batch_scene_with_non_finite_upload_is_refused():
    from basements_amd import Batch, BMError, SceneDesc

    P = random_scene(43, dx=DX)
    bad = {k: np.array(v, copy=True) for k, v in P.items()}
    bad["C"][0, 1, 1] = np.inf
    B = Batch(max_particles=100_000, max_blocks=20_000, max_scenes=4)
    sd = SceneDesc(dx=DX, origin=ORG, dims=DIMS)
    with pytest.raises(BMError) as ei:
        B.add_scene(sd, bad)
    assert ei.value.code == 4
    assert B.usage()[0] == 0  # nothing was added
    sid = B.add_scene(sd, P)
    B.step(1e-4, 3)
    assert B.scene_error(sid) == 0
