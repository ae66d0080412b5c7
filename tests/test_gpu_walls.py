"""Domain walls of the grid update (bm_mpm_set_walls, bm_scene_desc.walls).

The reference has only the floor (mpm_solver.h:263-288); the walls are this
library's addition, so they are pinned by their closed form: in a layer W
nodes deep on each face of the node box, a velocity component pointing out
of the box is zeroed after gravity; every other component -- inward motion,
the tangential part, everything outside the layer -- is left as the
reference's update leaves it, and nodes at or below the reference's mass
threshold (1e-6) keep their momentum.

Set-up: single particles with F = I and C = 0 (so P2G deposits w m v and no
stress), no gravity, a node box whose lowest node row sits above the floor
band (the floor rule then never applies), each particle near one face with
either an outward or an inward normal velocity.  The expected node velocity
is formed from the P2G grid the device downloaded (momentum / mass in FP32,
as the update divides), then the rule: bit-exact in both modes.  A batch
scene with walls is checked at particle level (one whole substep) against
the same closed form, within the fixed-point mass rounding (1e-4).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DX = 1.0 / 32.0
ORG = (0, 8, 0)    # node rows 8..31: (j + 1/2) dx >= dx everywhere, so no floor
DIMS = (24, 24, 24)
W = 3
DT = 1e-4


def _particles():
    """One particle per (face, direction).  Positions put the fraction at 1/2
    (weights 1/8, 3/4, 1/8) so that every node of the stencil carries mass well
    above 1e-6; particles sit at least 6 nodes apart so stencils never share a
    node."""
    from basements_amd.scenes import make_particles

    lo = [o + 2.0 for o in ORG]            # stencil base o+1: nodes o+1 .. o+3 (two in the layer)
    hi = [o + n - 3.0 for o, n in zip(ORG, DIMS)]  # base o+n-4: nodes o+n-4 .. o+n-2 (two in the layer)
    mid = [o + n / 2.0 for o, n in zip(ORG, DIMS)]
    xs, vs = [], []
    rng = np.random.default_rng(5)
    lanes = [-6.0, 0.0, 6.0]
    for axis in range(3):
        others = [a for a in range(3) if a != axis]
        for side, edge in (("lo", lo), ("hi", hi)):
            for k, sign in enumerate((-1.0, 1.0)):  # outward and inward
                p = list(mid)
                p[axis] = edge[axis]
                p[others[0]] = mid[others[0]] + lanes[k + (0 if side == "lo" else 1)]
                p[others[1]] = mid[others[1]] + (3.0 if side == "lo" else -3.0)
                xs.append([c * DX for c in p])
                v = rng.uniform(-0.5, 0.5, size=3)
                outward = -1.0 if side == "lo" else 1.0
                v[axis] = sign * outward * rng.uniform(0.3, 0.8)
                vs.append(v)
    return make_particles(np.array(xs), v=np.array(vs), mass=1e-2, rho=1600.0)


def _solver(mode, P):
    from basements_amd import Solver

    S = Solver(DX, ORG, DIMS, P["x"].shape[0], mode)
    S.set_material(3.5e5, 0.3, 1600.0)
    S.set_gravity((0.0, 0.0, 0.0))
    S.set_walls(W)
    S.upload(P)
    return S


def _wall_rule(vel, mass):
    """The closed form on a dense node box: zero outward components in the layer."""
    out = vel.copy()
    n = np.array(DIMS)
    idx = np.indices(DIMS)
    upd = mass > np.float32(1e-6)
    for a in range(3):
        low = (idx[a] < W) & (out[..., a] < 0.0) & upd
        high = (idx[a] >= n[a] - W) & (out[..., a] > 0.0) & upd
        out[..., a][low | high] = 0.0
    return out


@pytest.mark.parametrize("mode", [0, 1], ids=["reference", "fast"])
def test_walls_zero_outward_velocity_in_the_layer(mode):
    P = _particles()
    S = _solver(mode, P)
    q_exp = S.mass_quantum_exp()
    S.p2g(DT)
    G0 = S.download_grid()
    S.grid_update(DT)
    G1 = S.download_grid()
    mass = (G0["mass_q"].astype(np.float64) * 2.0 ** q_exp).astype(np.float32)
    act = G0["mass_q"] != 0
    assert act.sum() == 27 * P["x"].shape[0], "stencils overlap or left the box"
    mom = G0["vel"].astype(np.float32)
    if mode == 0:  # reference mode keeps the stress force apart (zero for F = I)
        assert not np.any(G0["force"][act])
    v = np.where((mass > np.float32(1e-6))[..., None], mom / np.where(mass > 0, mass, 1)[..., None], mom)
    want = _wall_rule(v.astype(np.float32), mass)
    got = G1["vel"]
    np.testing.assert_array_equal(got[act], want[act])
    # the rule did something on every face, and left inward motion alone
    idx = np.indices(DIMS)
    for a in range(3):
        layer = act & ((idx[a] < W) | (idx[a] >= DIMS[a] - W))
        zeroed = layer & (v[..., a] != 0.0) & (got[..., a] == 0.0)
        kept = layer & (v[..., a] != 0.0) & (got[..., a] == v[..., a])
        assert zeroed.any() and kept.any(), a
    # nodes outside every layer are untouched
    inner = act.copy()
    for a in range(3):
        inner &= (idx[a] >= W) & (idx[a] < DIMS[a] - W)
    np.testing.assert_array_equal(got[inner], v[inner])


def _g2p_closed_form(P, dx, org, dims):
    """One substep of isolated F = I particles with no gravity, closed form:
    node velocity = the particle's velocity (up to fixed-point mass
    rounding), the wall rule per node, then the particle gathers
    v' = sum w rule(v) and moves x' = x + v' dt."""
    x = P["x"].astype(np.float64)
    v = P["v"].astype(np.float64)
    n = np.array(dims)
    vn, xn = [], []
    for p in range(x.shape[0]):
        f = x[p] / dx - 0.5
        b = np.floor(f).astype(int)
        fr = f - b
        w = np.stack([0.5 * (1 - fr) ** 2, 0.75 - (fr - 0.5) ** 2, 0.5 * fr ** 2], axis=1)  # (axis, k)
        acc = np.zeros(3)
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    wt = w[0, i] * w[1, j] * w[2, k]
                    node = np.array([b[0] + i, b[1] + j, b[2] + k]) - np.array(org)
                    vv = v[p].copy()
                    for a in range(3):
                        if (node[a] < W and vv[a] < 0) or (node[a] >= n[a] - W and vv[a] > 0):
                            vv[a] = 0.0
                    acc += wt * vv
        vn.append(acc)
        xn.append(x[p] + acc * DT)
    return np.array(xn), np.array(vn)


@pytest.mark.parametrize("mode", [0, 1], ids=["reference", "fast"])
def test_walls_whole_substep_single_handle(mode):
    P = _particles()
    S = _solver(mode, P)
    S.step(DT)
    got = S.download()
    xn, vn = _g2p_closed_form(P, DX, ORG, S.dims)
    np.testing.assert_allclose(got["v"], vn, rtol=0, atol=1e-4 * np.abs(P["v"]).max())
    np.testing.assert_allclose(got["x"], xn, rtol=0, atol=1e-7)


def test_walls_in_a_batch_scene():
    """The same particles as the second scene of a batch (the first scene has
    no walls and other constants), one substep: the wall rule of the scene's
    own descriptor, at particle level."""
    from basements_amd import Batch, SceneDesc
    from basements_amd.scenes import soil_block

    P = _particles()
    B = Batch(max_particles=20_000, max_blocks=4_000, max_scenes=4)
    other = soil_block(1 / 24, (4, 0, 4), (8, 4, 8), seed=2, rho=1600.0)
    B.add_scene(SceneDesc(dx=1 / 24, origin=(-4, -4, -4), dims=(16, 16, 16), E=2e5, rho=1600.0), other)
    sid = B.add_scene(SceneDesc(dx=DX, origin=ORG, dims=DIMS, E=3.5e5, rho=1600.0, gravity=(0.0, 0.0, 0.0), walls=W),
                      P)
    B.step(DT)
    got = B.download_scene(sid)
    xn, vn = _g2p_closed_form(P, DX, ORG, tuple((d + 3) // 4 * 4 for d in DIMS))
    np.testing.assert_allclose(got["v"], vn, rtol=0, atol=1e-4 * np.abs(P["v"]).max())
    np.testing.assert_allclose(got["x"], xn, rtol=0, atol=1e-7)
    # without walls the outward components survive: the rule is what the test sees
    assert np.abs(vn - P["v"]).max() > 0.1
