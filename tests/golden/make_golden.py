"""Write the golden fixtures of tests/golden/ from the reference itself.

Runs only where the reference tree is present: it drives oracle/_ref/libref_mpm.so
(the reference's MPMSolver, SVD::compute, M0, M1 and TimeIntegrator compiled
unchanged from their sources by oracle/Makefile) on small seeded inputs and
stores inputs and outputs as .npz (no pickles).  It also copies the numbers of
the reference's committed repose-pile run
(research/mpm_rigid_coupling/outputs/csv/repose_pile*.csv) into a JSON data
fixture.  tests/test_golden.py checks the oracle against these files, so the
oracle stays pinned where the reference tree is absent.

    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import csv
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF_CSV = "/root/reference/research/mpm_rigid_coupling/outputs/csv"

from oracle.pyoracle import Body, Collider, Params, RefSim, make_particles, ref_available, round32  # noqa: E402

# Shared with tests/test_golden.py: the inputs are rebuilt there from these seeds.
SCENE = dict(n=160, lo=0.20, hi=0.35, dx=1.0 / 64.0, dt=1e-4)
PARAMS = dict(E=3.5e5, nu=0.3, rho=1600.0, alpha=0.386, mu_floor=0.4)
BOX_ORG, BOX_DIMS = (10, 10, 10), (16, 16, 16)


def scene(seed):
    rng = np.random.default_rng(seed)
    n = SCENE["n"]
    x = rng.uniform(SCENE["lo"], SCENE["hi"], size=(n, 3))
    P = round32(make_particles(x, 0.1 * rng.normal(size=(n, 3)), mass=7.6e-4, rho=PARAMS["rho"]))
    P["F"] = np.asarray(np.eye(3) + 0.03 * rng.normal(size=(n, 3, 3)), np.float32).astype(np.float64)
    P["C"] = np.asarray(0.5 * rng.normal(size=(n, 3, 3)), np.float32).astype(np.float64)
    return P


def svd_inputs(seed=7, n=96):
    rng = np.random.default_rng(seed)
    out = [np.eye(3), np.diag([1.5, 0.5, 1.5]), np.diag([1.0, 0.9, 1.0]), 0.95 * np.eye(3), np.zeros((3, 3)),
           np.diag([1e-7, 1.0, 2.0]), np.array([[1, 1e-7, 0], [0, 1, 0], [0, 0, 1.0]]), np.diag([1.2, 1.2, 0.8])]
    for k in range(n):
        if k % 3 == 0:
            out.append(np.eye(3) + 1e-3 * rng.normal(size=(3, 3)))
        elif k % 3 == 1:
            out.append(rng.normal(size=(3, 3)))
        else:
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            out.append(q @ np.diag(rng.uniform(0.5, 1.5, 3)) @ q.T)
    return np.asarray(out, np.float32).astype(np.float64)


def colliders():
    q = np.array([0.9, 0.1, 0.3, 0.2])
    q = tuple(np.asarray(q / np.linalg.norm(q), np.float32).astype(np.float64))
    return [Collider(center=(0.28, 0.3, 0.27), quat=q, half_extents=(0.04, 0.05, 0.045), lin_vel=(0.1, -0.5, 0.2),
                     ang_vel=(1.0, 2.0, 3.0), friction=0.5, mass=2.0, shape=s) for s in (0, 1)]


def body0():
    return Body(x=(0.1, 0.5, -0.2), quat=(0.9, 0.1, 0.3, 0.2), v=(0.3, -1.0, 0.2), w=(1.0, -2.0, 0.5), inv_mass=0.5,
                inv_I=np.diag([2.0, 3.0, 4.0]), lin_damp=0.01, ang_damp=0.02, force=(1.0, 2.0, -3.0),
                torque=(0.1, 0.2, 0.3))


def _f32(a):
    a = np.asarray(a, np.float64)
    assert np.array_equal(a, a.astype(np.float32).astype(np.float64)), "reference outputs are FP32"
    return a.astype(np.float32)


def _grid(sim):
    g = sim.grid_dense(BOX_ORG, BOX_DIMS)
    return {"mass": _f32(g["mass"]), "vel": _f32(g["vel"]), "force": _f32(g["force"]),
            "active": g["active"].astype(np.uint8), "alloc": g["alloc"].astype(np.uint8)}


def _particles(P):
    return {k: _f32(P[k]) for k in "xvFC"}


def main():
    if not ref_available():
        raise SystemExit("oracle/_ref/libref_mpm.so is not built (run make -C oracle where the reference is present)")
    prm = Params(**PARAMS)
    dx, dt = SCENE["dx"], SCENE["dt"]

    # SVD::compute
    Fs = svd_inputs()
    r = RefSim(dx, params=prm)
    U, S, V = zip(*[r.svd(F) for F in Fs])
    np.savez_compressed(os.path.join(HERE, "ref_svd.npz"), F=_f32(Fs), U=_f32(U), S=_f32(S), V=_f32(V))

    # one substep, phase by phase
    r = RefSim(dx, params=prm)
    r.set_particles(scene(1))
    r.p2g()
    out = {f"p2g_{k}": v for k, v in _grid(r).items()}
    r.update(dt)
    out.update({f"update_{k}": v for k, v in _grid(r).items()})
    r.g2p(dt)
    out.update({f"g2p_{k}": v for k, v in _particles(r.get_particles()).items()})
    np.savez_compressed(os.path.join(HERE, "ref_phases.npz"), **out)

    # M0 and M1 on a particle-driven grid, box and cylinder
    out = {}
    for method in (0, 1):
        for col in colliders():
            r = RefSim(dx, params=prm)
            r.set_particles(scene(2))
            r.p2g()
            r.update(dt)
            f, t, n = r.couple(method, col, dt)
            tag = f"m{method}_s{col.shape}"
            out[f"{tag}_reaction"] = _f32(np.concatenate([f, t]))
            out[f"{tag}_nodes"] = np.array([n], np.int32)
            out[f"{tag}_vel"] = _grid(r)["vel"]
            out[f"{tag}_mass"] = _grid(r)["mass"]
    np.savez_compressed(os.path.join(HERE, "ref_couplers.npz"), **out)

    # 50 whole substeps (MPMSolver::step)
    r = RefSim(dx, params=prm)
    r.set_particles(scene(3))
    r.step(dt, 50)
    np.savez_compressed(os.path.join(HERE, "ref_trajectory.npz"), **_particles(r.get_particles()))

    # TimeIntegrator::integrate, 20 steps at 1/240
    r = RefSim(dx, params=prm)
    b, seq = body0(), []
    for _ in range(20):
        b2 = r.integrate_body(b, 1.0 / 240.0)
        seq.append(b2.pack())
        b = Body(x=b2.x, quat=b2.quat, v=b2.v, w=b2.w, inv_mass=b.inv_mass, inv_I=b.inv_I, lin_damp=b.lin_damp,
                 ang_damp=b.ang_damp, force=(1.0, 2.0, -3.0), torque=(0.1, 0.2, 0.3))
    np.savez_compressed(os.path.join(HERE, "ref_body.npz"), seq=np.asarray(seq, np.float64))

    # the reference's committed repose-pile outputs (data, as printed)
    with open(os.path.join(REF_CSV, "repose_pile_summary.csv")) as f:
        summ = next(csv.DictReader(f))
    with open(os.path.join(REF_CSV, "repose_pile.csv")) as f:
        series = [[float(x) for x in row] for row in csv.reader(f) if row and row[0] != "t"]
    data = {
        "source": "research/mpm_rigid_coupling/outputs/csv/repose_pile_summary.csv and repose_pile.csv",
        "scenario": {"dx": 0.025, "column": [7, 5, 7], "spacing": 0.0125, "mass": 2.93e-3, "base": [0.0, 0.0125, 0.0],
                     "dt": 1.0 / 960.0, "steps": 1920, "E": 5.0e5, "nu": 0.3, "rho": 1500.0, "alpha": 0.55,
                     "mu_floor": 0.6, "gravity": 9.81, "settle_time": 1.5, "bins": 24},
        "summary": {k: (float(v) if k != "literature_target_deg" else v) for k, v in summ.items()},
        "series_columns": ["t", "height", "radius", "angle_deg", "particles_at_rest"],
        "series": series,
    }
    with open(os.path.join(HERE, "repose_pile.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
