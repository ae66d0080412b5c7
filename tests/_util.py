"""Shared helpers for the parity tests: identical inputs for the HIP path and the oracle."""
from __future__ import annotations

import numpy as np

from oracle.pyoracle import MASS_FIXED, Collider as OCollider, OracleSim, Params


def random_scene(seed, dx=1.0 / 32.0, cells_lo=(6, 1, 6), cells_hi=(16, 7, 16), rho=1600.0, strain=0.05, vel=0.3,
                 cgrad=2.0):
    """Jittered soil block with random velocity, F and C (float32 values)."""
    from basements_amd.scenes import soil_block

    P = soil_block(dx, cells_lo, cells_hi, seed=seed, rho=rho)
    rng = np.random.default_rng(seed + 1000)
    n = P["x"].shape[0]
    P["v"] = (vel * rng.normal(size=(n, 3))).astype(np.float32)
    P["F"] = (np.eye(3) + strain * rng.normal(size=(n, 3, 3))).astype(np.float32)
    P["C"] = (cgrad * rng.normal(size=(n, 3, 3))).astype(np.float32)
    return P


def params_from(E=3.5e5, nu=0.3, rho=1600.0, alpha=0.386, mu_floor=0.4, g=(0.0, -9.81, 0.0)):
    return Params(E=E, nu=nu, rho=rho, alpha=alpha, mu_floor=mu_floor, gravity=g)


def gpu_solver(P, dx, origin, dims, prm: Params, mode=0, cap=None):
    from basements_amd import Solver

    s = Solver(dx, origin, dims, cap or max(1, P["x"].shape[0]), mode)
    s.set_material(prm.E, prm.nu, prm.rho)
    s.set_plastic_alpha(prm.alpha)
    s.set_floor_friction(prm.mu_floor)
    s.set_gravity(prm.gravity)
    s.upload(P)
    return s


def oracle_for(P, dx, prm: Params, q_exp, precision="f32"):
    o = OracleSim(dx, precision=precision, mass_mode=MASS_FIXED, q_exp=q_exp, params=prm)
    o.set_particles({k: np.asarray(v, np.float64) for k, v in P.items()})
    return o


def ocollider(col):
    return OCollider(center=col.center, quat=col.quat, half_extents=col.half_extents, lin_vel=col.lin_vel,
                     ang_vel=col.ang_vel, friction=col.friction, mass=col.mass, shape=col.shape)


def assert_close_field(got, want, rtol, what, scale=None):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    sc = float(np.abs(want).max()) if scale is None else scale
    err = np.abs(got - want)
    bound = rtol * np.maximum(np.abs(want), sc)
    bad = err > bound
    assert not bad.any(), (f"{what}: {int(bad.sum())} of {bad.size} entries off; worst rel "
                           f"{float((err / np.maximum(np.abs(want), sc + 1e-300)).max()):.3e}")


def rel_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))


def _r32(t):
    """Values as the device holds them (FP32), handed to the oracle in double."""
    return tuple(float(np.float32(v)) for v in np.ravel(t))


def oracle_coupled_run(P, dx, prm: Params, body, method, dt, n, precision):
    """The oracle's coupled loop (contract order, OracleSim.step_coupled) from
    the same particles and device body description (basements_amd.Body, its
    values rounded to FP32 as the device holds them).  Returns (particles,
    per-substep record)."""
    from oracle.pyoracle import Body as OBody

    if body.integrate not in (1, 2):
        raise ValueError("the oracle's coupled loop integrates dynamic or kinematic bodies")
    col = OCollider(half_extents=_r32(body.half_extents), friction=_r32([body.friction])[0],
                    mass=_r32([body.mass])[0], shape=int(body.shape))
    ob = OBody(x=_r32(body.position), quat=_r32(body.orientation), v=_r32(body.linear_velocity),
               w=_r32(body.angular_velocity), inv_mass=_r32([body.inv_mass])[0],
               inv_I=np.array(_r32(body.inv_inertia)).reshape(3, 3), lin_damp=_r32([body.linear_damping])[0],
               ang_damp=_r32([body.angular_damping])[0], force=_r32(body.force), torque=_r32(body.torque))
    o = OracleSim(dx, precision=precision, params=prm)
    o.set_particles({k: np.asarray(v, np.float64) for k, v in P.items()})
    _, rec = o.step_coupled(method, col, ob, _r32(body.applied_force), _r32(body.applied_torque), dt, n,
                            kinematic=body.integrate == 2)
    return o.get_particles(), rec


def oracle_pairs(jobs, workers=16):
    """Run oracle_coupled_run for every job (a dict of its keyword arguments
    without `precision`) in FP32 and FP64, in parallel threads (the oracle's
    calls release the interpreter lock).  Returns {(job index, precision): result}."""
    from concurrent.futures import ThreadPoolExecutor

    out = {}
    with ThreadPoolExecutor(max_workers=workers) as ex:
        futs = {(i, p): ex.submit(oracle_coupled_run, precision=p, **job) for i, job in enumerate(jobs)
                for p in ("f32", "f64")}
        for k, f in futs.items():
            out[k] = f.result()
    return out


def assert_at_noise_floor(got, P32, P64, what, fields="xvFC", factor=2.5, slack=1e-6):
    """Device vs the FP64 oracle within `factor` x the FP32 oracle's own
    distance from it (plus `slack`), per field relative to its largest
    magnitude.  Returns {field: (device error, FP32 floor)}."""
    rep = {}
    for k in fields:
        floor = rel_err(P32[k], P64[k])
        err = rel_err(got[k], P64[k])
        rep[k] = (err, floor)
        assert err <= factor * floor + slack, (what, k, err, floor)
    return rep
