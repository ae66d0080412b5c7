"""bm_mpm_couple_noalloc (the allocation-free coupler) against bm_mpm_couple.

Same kernels, the partial reactions in the library's fixed device array
instead of the handle's buffers, the reaction written to the caller's device
memory: the reaction must be bit-identical to bm_mpm_couple's and the grid
must be updated identically, on two calls in a row -- the second on a
smaller grid with a smaller body, after the first filled the shared array --
for M0 and M1; dt <= 0 writes a zero reaction.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DX = 1.0 / 32.0


def _grid(dims, origin, seed):
    """A seeded post-update grid: a slab of massive nodes with random
    velocities, empty nodes around it (M1 claims those)."""
    rng = np.random.default_rng(seed)
    dims = np.array(dims)
    mass = np.zeros(dims, np.float32)
    vel = np.zeros(tuple(dims) + (3,), np.float32)
    act = np.zeros(dims, np.uint8)
    lo = dims // 4
    hi = lo + dims // 2
    sl = tuple(slice(a, b) for a, b in zip(lo, hi))
    mass[sl] = rng.uniform(1e-4, 2e-3, size=tuple(hi - lo)).astype(np.float32)
    vel[sl] = rng.normal(0.0, 0.3, size=tuple(hi - lo) + (3,)).astype(np.float32)
    act[sl] = 1
    return mass, vel, act


def _pair(dims, origin, seed):
    from basements_amd import Solver

    out = []
    for _ in range(2):
        S = Solver(DX, origin, dims, 16, 0)
        S.upload_grid(origin, *_grid(dims, origin, seed))
        out.append(S)
    return out


def _read_reaction(t):
    import torch

    raw = t.cpu().numpy().tobytes()
    f = np.frombuffer(raw[:24], np.float32)
    n = int(np.frombuffer(raw[24:28], np.int32)[0])
    return f[:3].copy(), f[3:6].copy(), n


def _device_reaction():
    import torch

    t = torch.full((28,), 0xFF, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    return t


def _collider(method, center, half):
    from basements_amd import SHAPE_BOX, SHAPE_CYLINDER, Collider

    q = (math.cos(0.3), math.sin(0.3) * 0.6, 0.0, math.sin(0.3) * 0.8)
    return Collider(center=center, quat=q, half_extents=half, lin_vel=(0.1, -0.4, 0.05), ang_vel=(0.5, -1.0, 0.2),
                    friction=0.5, mass=2.0, shape=SHAPE_BOX if method == 0 else SHAPE_CYLINDER)


@pytest.mark.parametrize("method", [0, 1], ids=["M0", "M1"])
def test_noalloc_coupler_equals_couple_twice(method):
    dt = 2e-4
    # first call: a 40 x 32 x 40 box and a large body
    big = ((-4, -4, -4), (40, 32, 40), (0.5, 0.35, 0.5), (0.12, 0.08, 0.1))
    # second call: a smaller 24^3 box and a smaller body
    small = ((-4, -4, -4), (24, 24, 24), (0.25, 0.2, 0.25), (0.05, 0.04, 0.06))
    for k, (origin, dims, center, half) in enumerate((big, small)):
        A, B = _pair(dims, origin, seed=10 * method + k)
        col = _collider(method, center, half)
        f, t, n = A.couple(method, col, dt)
        assert n > 0, "the body touched no node: the test would pin nothing"
        r = _device_reaction()
        B.couple_noalloc(method, col, dt, r.data_ptr())
        B.synchronize()
        gf, gt, gn = _read_reaction(r)
        np.testing.assert_array_equal(gf, np.asarray(f, np.float32))
        np.testing.assert_array_equal(gt, np.asarray(t, np.float32))
        assert gn == n
        ga, gb = A.download_grid(), B.download_grid()
        for key in ("vel", "mass", "active"):
            np.testing.assert_array_equal(ga[key], gb[key], err_msg=f"call {k} {key}")


def test_noalloc_coupler_zero_dt_writes_a_zero_reaction():
    (S, _) = _pair((24, 24, 24), (-4, -4, -4), seed=3)
    r = _device_reaction()
    S.couple_noalloc(0, _collider(0, (0.25, 0.2, 0.25), (0.05, 0.05, 0.05)), 0.0, r.data_ptr())
    S.synchronize()
    f, t, n = _read_reaction(r)
    assert not f.any() and not t.any() and n == 0
