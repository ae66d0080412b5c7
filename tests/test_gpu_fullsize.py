"""The measured workload at its full size (4,147,200 particles on a 256^3
node box) and the smallest inputs.

Full size: one reference-order substep against the FP32 oracle on the same
input; twenty coupled fast-mode substeps with the rolling wheel through the
properties that do not depend on size -- every particle's mass quanta are on
the grid (bit-exact), the state stays finite and inside the box, the body
stays finite and the FP64 SVD fallback stays rare.  Smallest: no particles,
one particle, and ragged counts against the oracle."""
from __future__ import annotations

import math

import numpy as np
import pytest

from _util import assert_close_field, gpu_solver, oracle_for, params_from, rel_err

pytestmark = pytest.mark.gpu

DT = 1e-4


def _wheel_bed():
    from basements_amd.scenes import wheel_bed

    return wheel_bed(seed=0)


def _quanta_total(m, q_exp):
    return int(np.rint(np.asarray(m, np.float64) * 2.0 ** (-q_exp)).astype(np.int64).sum())


def test_full_size_reference_substep_matches_oracle():
    sc = _wheel_bed()
    P = sc.particles
    prm = params_from(E=sc.E, nu=sc.nu, rho=sc.rho, alpha=sc.alpha, mu_floor=sc.mu_floor, g=sc.gravity)
    S = gpu_solver(P, sc.dx, sc.origin, sc.dims, prm, mode=0)
    q_exp = S.mass_quantum_exp()
    S.p2g(DT)
    S.synchronize()
    assert S.grid_mass_total() == _quanta_total(P["m"], q_exp)
    S.grid_update(DT)
    S.g2p(DT)
    got = S.download()
    o = oracle_for(P, sc.dx, prm, q_exp)
    o.step(DT)
    want = o.get_particles()
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], want[key], 1e-4, f"full-size {key}")


def test_full_size_fast_coupled_run_keeps_mass_and_stays_finite():
    from basements_amd import M1, SHAPE_CYLINDER, Body

    sc = _wheel_bed()
    P = sc.particles
    prm = params_from(E=sc.E, nu=sc.nu, rho=sc.rho, alpha=sc.alpha, mu_floor=sc.mu_floor, g=sc.gravity)
    S = gpu_solver(P, sc.dx, sc.origin, sc.dims, prm, mode=1)
    t = sc.tool
    m = t["load"] / 9.81
    S.set_body(Body(position=t["center"], orientation=(math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0),
                    linear_velocity=(t["speed"], 0.0, 0.0), angular_velocity=(0.0, 0.0, -t["speed"] / t["radius"]),
                    inv_mass=1.0 / m, linear_damping=0.0, angular_damping=0.0, applied_force=(0.0, -9.81 * m, 0.0),
                    half_extents=(t["radius"], t["half_len"], t["radius"]), friction=t["friction"], mass=m,
                    shape=SHAPE_CYLINDER))
    want_q = _quanta_total(P["m"], S.mass_quantum_exp())
    for _ in range(4):
        S.step_coupled(M1, DT, 5)
        S.synchronize()
        assert S.grid_mass_total() == want_q  # the deposit for the next substep holds every particle
    got = S.download()
    for key in ("x", "v", "F", "C"):
        assert np.isfinite(got[key]).all(), key
    lo = (np.array(sc.origin) + 1) * sc.dx
    hi = (np.array(sc.origin) + np.array(sc.dims) - 2) * sc.dx
    assert (got["x"] >= lo).all() and (got["x"] <= hi).all()
    body, reaction = S.get_body()
    assert all(math.isfinite(v) for v in (*body.position, *body.linear_velocity))
    assert reaction[2] > 0  # the wheel rests on soil nodes
    assert S.stats()["fp64_fallbacks"] < 0.01 * P["x"].shape[0]


@pytest.mark.parametrize("mode", [0, 1])
def test_no_particles(mode):
    from basements_amd import Solver

    S = Solver(1 / 32, (-4, -4, -4), (24, 24, 24), 16, mode)
    S.step(DT, 3)
    S.synchronize()
    assert S.num_particles() == 0
    assert S.grid_mass_total() == 0
    assert S.download()["x"].shape == (0, 3)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", [1, 65, 257])
def test_small_ragged_counts_match_oracle(mode, n):
    """1, 65 and 257 particles (not multiples of a wave or a workgroup) in a
    loose random cloud, 20 substeps.  A loose cloud puts single particles at
    the edge of their nodes' support, where v = momentum / mass amplifies
    rounding, so the bar is the FP32 noise floor of the fast-mode tests:
    against the FP64 oracle, within 2.5x the FP32 oracle's own distance
    (plus 1e-6)."""
    from oracle.pyoracle import OracleSim
    from basements_amd.scenes import make_particles

    rng = np.random.default_rng(n)
    x = 0.25 + 0.2 * rng.random((n, 3))
    P = make_particles(x, v=0.1 * rng.normal(size=(n, 3)), mass=1e-3, rho=1600.0)
    prm = params_from()
    S = gpu_solver(P, 1 / 32, (-4, -4, -4), (32, 32, 32), prm, mode=mode)
    S.step(DT, 20)
    got = S.download()
    Pd = {k: np.asarray(v, np.float64) for k, v in P.items()}
    runs = {}
    for prec in ("f32", "f64"):
        o = OracleSim(1 / 32, prec, params=prm)
        o.set_particles(Pd)
        o.step(DT, 20)
        runs[prec] = o.get_particles()
    for key in ("x", "v", "F", "C"):
        floor = rel_err(runs["f32"][key], runs["f64"][key])
        err = rel_err(got[key], runs["f64"][key])
        assert err <= 2.5 * floor + 1e-6, (n, mode, key, err, floor)
