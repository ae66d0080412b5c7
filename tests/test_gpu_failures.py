"""Failure detection at the boundary: non-finite state is refused at upload
(BM_ERR_NONFINITE) and leaves the handle usable; a non-finite value that
reaches a particle on the device is flagged by the substep, not propagated
silently; a scene of a batch that goes non-finite fails alone."""
from __future__ import annotations

import numpy as np
import pytest

from _util import assert_close_field, gpu_solver, params_from, random_scene

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


def _edge_runner(n_cells=2, seed=44):
    """Particles whose stencil base is x-cell 13 of a 16-node box (valid: the
    stencil ends at node 15) moving at +94 m/s, i.e. 0.3 cells per 1e-4 s
    substep: the G2P moves them to base 14, whose stencil ends past the box."""
    from basements_amd.scenes import soil_block

    P = soil_block(DX, (13, 6, 6), (14, 6 + n_cells, 6 + n_cells), seed=seed, rho=1600.0)
    P["x"][:, 0] = np.float32(14.4 * DX)
    P["v"][:, 0] = np.float32(0.3 * DX / 1e-4)
    return P


def test_particle_moved_past_the_box_edge_is_flagged_not_deposited():
    """The deposit that follows a G2P must not write the nodes past the box
    for a particle the G2P just moved there (its stencil base is inside the
    unit's block, but its stencil is not inside the box): the substep flags
    BM_ERR_OUT_OF_DOMAIN at the next synchronisation."""
    from basements_amd import BMError, MODE_FAST

    P = _edge_runner()
    S = gpu_solver(P, DX, (0, 0, 0), (16, 16, 16), params_from(g=(0.0, 0.0, 0.0)), mode=MODE_FAST)
    S.step(1e-4, 1)
    with pytest.raises(BMError) as ei:
        S.synchronize()
    assert ei.value.code == 6


def test_scene_whose_particle_leaves_its_box_cannot_write_into_the_next_scene():
    """In a batch, the nodes just past scene A's box on +x are the first
    nodes of scene B's blocks.  A's particles moved past its edge are flagged
    and not deposited; B, whose soil sits in exactly those blocks, comes out
    as it does alone."""
    from basements_amd import Batch, SceneDesc

    A = _edge_runner()
    Bp = random_scene(45, dx=DX, cells_lo=(1, 8, 8), cells_hi=(5, 12, 12), vel=0.1, strain=0.01, cgrad=0.5)
    prm = params_from(g=(0.0, 0.0, 0.0))
    sd = SceneDesc(dx=DX, origin=(0, 0, 0), dims=(16, 16, 16), E=prm.E, nu=prm.nu, rho=prm.rho, alpha=prm.alpha,
                   mu_floor=prm.mu_floor, gravity=prm.gravity)

    def run(scenes):
        B = Batch(max_particles=20_000, max_blocks=1_000, max_scenes=4)
        ids = [B.add_scene(sd, p) for p in scenes]
        B.step(1e-4, 2)
        return B, ids

    Bm, (ia, ib) = run([A, Bp])
    assert Bm.scene_error(ia) == 6
    assert Bm.scene_error(ib) == 0
    got = Bm.download_scene(ib)
    Ba, (ib1,) = run([Bp])
    want = Ba.download_scene(ib1)
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], want[key], 1e-5, f"scene B {key}")
