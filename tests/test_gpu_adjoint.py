"""The differentiable rollout (bm_diff_*): its forward substep against the FP64
oracle, and its hand-written adjoint against FP64 central finite
differences of the same forward.

The reference has no backward pass (its tape records nothing), so finite
differences are the only pin for the gradient: per parameter, step
h = 1e-4 relative to the parameter (to the tool speed for velocity
components), and the adjoint must agree with (L(p+h) - L(p-h)) / 2h to 1 %
of the largest gradient magnitude of that kind -- the < 1 % L2 criterion of
the reference's differentiable-MPM plan, applied per parameter group.
The forward model is pinned to the oracle (the reference's algorithm
restated) at FP64, without and with the M0 tool."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from _util import assert_close_field

pytestmark = pytest.mark.gpu

SMALL = (12, 4, 12)  # 4,608 particles


def _scene(velocity=0.3, cells=SMALL):
    from basements_amd.scenes import tine_scene

    return tine_scene(seed=3, cells=cells, velocity=velocity)


def _rollout(sc, precision=64, K=16, reference_svd=False):
    from basements_amd import Rollout

    R = Rollout(sc["dx"], sc["origin"], sc["dims"], sc["particles"]["x"].shape[0], precision=precision,
                checkpoint_every=K, reference_svd=reference_svd)
    R.set_params(sc["params"])
    R.upload(sc["particles"], sc["target"])
    return R


def _loss(sc, prm, substeps, precision=64):
    from basements_amd import Rollout

    R = Rollout(sc["dx"], sc["origin"], sc["dims"], sc["particles"]["x"].shape[0], precision=precision)
    R.set_params(prm)
    R.upload(sc["particles"], sc["target"])
    return R.forward(sc["dt"], substeps)


def _fd(sc, substeps, name, idx=None, rel=1e-4):
    p0 = sc["params"]
    if name == "tool_velocity":
        h = rel * float(np.linalg.norm(p0.tool_velocity))
        v = list(p0.tool_velocity)
        vp, vm = list(v), list(v)
        vp[idx] += h
        vm[idx] -= h
        pp, pm = dataclasses.replace(p0, tool_velocity=tuple(vp)), dataclasses.replace(p0, tool_velocity=tuple(vm))
    else:
        val = getattr(p0, name)
        h = rel * abs(val)
        pp, pm = dataclasses.replace(p0, **{name: val + h}), dataclasses.replace(p0, **{name: val - h})
    return (_loss(sc, pp, substeps) - _loss(sc, pm, substeps)) / (2 * h)


@pytest.mark.parametrize("substeps", [40, 48])  # a partial last segment, whole segments
def test_adjoint_matches_central_differences(substeps):
    sc = _scene()
    R = _rollout(sc)
    loss = R.forward(sc["dt"], substeps)
    assert loss > 0
    g = R.backward()
    for name, val, fd in (("young_modulus", g["young_modulus"], _fd(sc, substeps, "young_modulus")),
                          ("friction_deg", g["friction_deg"], _fd(sc, substeps, "friction_deg"))):
        assert val == pytest.approx(fd, rel=1e-2), (name, val, fd)
    fdv = np.array([_fd(sc, substeps, "tool_velocity", a) for a in range(3)])
    scale = float(np.abs(fdv).max())
    assert scale > 0  # the tine moves soil
    assert np.abs(g["tool_velocity"] - fdv).max() <= 1e-2 * scale, (g["tool_velocity"], fdv)


def test_checkpoint_interval_does_not_change_the_gradient():
    """K = 1, 16, 48, each with the substeps' grids kept by the forward (the
    default budget) and recomputed per segment by the backward (budget 0)."""
    sc = _scene()
    out = []
    for K in (1, 16, 48):
        for budget in (None, 0):
            R = _rollout(sc, K=K)
            if budget is not None:
                R.set_grid_budget(budget)
            R.forward(sc["dt"], 48)
            out.append(R.backward())
    for g in out[1:]:
        for key in ("young_modulus", "friction_deg"):
            assert g[key] == pytest.approx(out[0][key], rel=1e-9)
        assert np.allclose(g["tool_velocity"], out[0]["tool_velocity"], rtol=1e-9, atol=1e-15)


def test_fp32_adjoint_tracks_fp64():
    sc = _scene()
    g = {}
    for prec in (32, 64):
        R = _rollout(sc, precision=prec)
        R.forward(sc["dt"], 48)
        g[prec] = R.backward()
    for key in ("young_modulus", "friction_deg"):
        assert g[32][key] == pytest.approx(g[64][key], rel=5e-2), key
    sc64 = float(np.abs(g[64]["tool_velocity"]).max())
    assert np.abs(g[32]["tool_velocity"] - g[64]["tool_velocity"]).max() <= 5e-2 * sc64


def _oracle_run(sc, substeps, with_tool):
    from oracle.pyoracle import MASS_FLOAT, Collider, OracleSim, Params
    from basements_amd.scenes import dp_alpha

    p = sc["params"]
    prm = Params(E=p.young_modulus, nu=p.poisson_ratio, rho=1600.0, alpha=dp_alpha(p.friction_deg),
                 mu_floor=p.floor_friction, gravity=p.gravity)
    o = OracleSim(sc["dx"], precision="f64", mass_mode=MASS_FLOAT, params=prm)
    o.set_particles({k: np.asarray(v, np.float64) for k, v in sc["particles"].items()})
    for s in range(substeps):
        o.p2g()
        o.update(sc["dt"])
        if with_tool:
            c = tuple(p.tool_center[a] + p.tool_velocity[a] * s * sc["dt"] for a in range(3))
            o.couple(0, Collider(center=c, quat=p.tool_quat, half_extents=p.tool_half_extents,
                                 lin_vel=p.tool_velocity, ang_vel=(0.0, 0.0, 0.0), friction=p.tool_friction,
                                 mass=1.0, shape=0), sc["dt"])
        o.g2p(sc["dt"])
    return o.get_particles()


@pytest.mark.parametrize("with_tool", [False, True])
def test_forward_matches_the_fp64_oracle(with_tool):
    """The rollout's forward substeps are the reference algorithm: with the
    reference's SVD stopping rule, against the FP64 oracle (float mass) over
    24 substeps to 1e-6 of each field's scale (the state comes back in FP32).
    With the tighter SVD the adjoint uses, the rollout differs from the
    reference by what the reference's Jacobi leaves unconverged (pairs with
    |(F^T F)_ij| < 1e-6 are never rotated): a few 1e-5 here."""
    sc = _scene()
    if not with_tool:
        sc["params"] = dataclasses.replace(sc["params"], tool=False)
    want = _oracle_run(sc, 24, with_tool)
    R = _rollout(sc, reference_svd=True)
    R.forward(sc["dt"], 24)
    got = R.download()
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], want[key], 1e-6, f"{key} (tool={with_tool})")
    R = _rollout(sc)
    R.forward(sc["dt"], 24)
    got = R.download()
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], want[key], 2e-4, f"{key} (tool={with_tool}, tight SVD)")
