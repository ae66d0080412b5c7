#!/usr/bin/env python3
"""Throughput of the MLS-MPM soil substep on one MI355X.

Workload (default): the single large scene -- a 1.0 x 0.3 x 1.0 m soil bed at
8 particles per cell (dx = 1/120 m, 4,147,200 particles) on a 256^3-node grid,
friction angle from U(25, 40) deg, a cylinder wheel (r 0.1 m, width 0.08 m)
rolling at 0.2 m/s under a 50 N load through the M1 PIC-rigid coupler, dt =
1e-4 s.  One step = one coupled substep: P2G, grid update, M1 coupler with the
reaction added to the wheel (a dynamic body under its 50 N weight), wheel
integration, G2P with plasticity -- all on the device, no host round trip.

Prints ONE JSON line: particle-updates/s (value), the roofline of the
dominant kernel (HIP events on the handle's stream), and a CPU baseline (the
reference's own MPMSolver::step built from its sources, one thread, on a
bounded sample of the same bed).  Beside them the same line carries the
batched and served forms: `batched` -- 64 uneven tool scenes (2e4-5e5
particles) stepped together against the same scenes one at a time on the
plain path; `served` -- 48 soil-tool requests (5e4-2e6 particles, 200-2000
substeps) arriving at 200 requests/s through the request loop, with p50 /
p99 latency, against the same requests one at a time.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes of each timed phase, (per particle, per active node): the
# inputs and outputs of the reference operation the phase completes.
#   P2G (p2g_single) reads x v C F m V0 (104 B) and adds m, momentum and
#   force into each active node (28 B); the update reads those 28 B and
#   writes v (12 B); G2P (g2p_single + apply_plasticity) reads x F (48 B) and
#   v of each active node (12 B) and writes x v C F (96 B).
# The substep as a whole must move 200 B per particle (state read once,
# written once) + 80 B per active node.
BYTES_REF = {
    "p2g": (104, 28),
    "update": (0, 40),
    "g2p": (144, 12),
}
# Fast mode folds the stress force into the momentum (56 B per active node
# per substep instead of 80): P2G deposits m + momentum (16 B), the update
# reads 16 B and writes 12 B, G2P reads 12 B.  Two-kernel form: the particle
# pass (phase g2p) completes G2P; the cell deposit (phase p2g) completes the
# next substep's P2G.
BYTES_FAST = {
    "p2g": (104, 16),
    "update": (0, 28),
    "g2p": (144, 12),
}
# Kernel that each phase's roofline describes (names as rocprofv3 lists them).
KERNEL_OF = {
    ("fast", "g2p"): "k_advance_tile",
    ("fast", "p2g"): "k_p2g_cell<1>",
    ("fast", "update"): "k_update",
    ("reference", "p2g"): "k_p2g_ref",
    ("reference", "update"): "k_update",
    ("reference", "g2p"): "k_g2p_ref",
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
REBIN_PERIOD = 32   # a single handle's default re-binning period (substeps)
BATCH_STEPS = 64    # the default line's batched measurement: two re-binning cycles of 64 scenes
ADJ_BYTES_PER_PARTICLE_SUBSTEP = 600  # adjoint rollout: forward 200 B + backward 400 B per particle and substep
SERVE_RATE = 200.0  # the default line's served measurement: a saturating arrival rate (requests/s)


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="wheel_bed",
                    choices=["wheel_bed", "box_drop", "batch", "serve", "adjoint", "coupler_loops"],
                    help="wheel_bed (the default workload), box_drop, or batch: 64 uneven tool scenes stepped "
                         "together against the same scenes one at a time on the single-scene path; "
                         "coupler_loops: the research harness's grid-only M0 / M1 loops")
    ap.add_argument("--scenes", type=int, default=64, help="scenes of --config batch")
    ap.add_argument("--graph", type=int, default=1, help="--config batch: replay captured substeps")
    ap.add_argument("--serial", type=int, default=1, help="--config batch/serve: also time the work one at a time")
    ap.add_argument("--requests", type=int, default=48, help="--config serve: requests per arrival rate")
    ap.add_argument("--fd-check", type=int, default=1, help="--config adjoint: FP64 central-difference check")
    ap.add_argument("--rollouts", type=int, default=8,
                    help="--config adjoint: rollouts of a parameter sweep run side by side in one handle (0 = skip)")
    ap.add_argument("--single-rollout", type=int, default=1,
                    help="--config adjoint: 0 = time only the --rollouts sweep (a profile of the batched form)")
    ap.add_argument("--rates", default="5,200", help="--config serve: arrival rates (requests/s)")
    ap.add_argument("--serve-particles", type=int, default=16_000_000, help="--config serve: particle room")
    ap.add_argument("--serve-blocks", type=int, default=400_000, help="--config serve: grid-block room")
    ap.add_argument("--tuning", default="",
                    help="--config batch: unit,sort_every,deposit_groups instead of autotune; "
                         "wheel_bed / box_drop fast mode: unit,sort_every")
    ap.add_argument("--mode", default="fast", choices=["reference", "fast"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--batched", type=int, default=1,
                    help="default workload: also measure 64 batched scenes against the same scenes one at a time")
    ap.add_argument("--served", type=int, default=1,
                    help="default workload: also measure 48 served requests at a saturating arrival rate")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (for HBM counter scaling)")
    return ap.parse_args()


def copy_calibration():
    import torch

    n = 256 << 20  # 256 Mi floats = 1 GiB each way
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)  # the runtime's copy kernel
    for _ in range(5):
        torch.mul(a, 1.0, out=b)  # an elementwise kernel: vector loads and stores, as the solver's kernels
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 5}))


def make_body(tool, g=9.81):
    """The scene's dynamic body (basements_amd.scenes.dynamic_body)."""
    from basements_amd.scenes import dynamic_body

    return dynamic_body(tool, g)


def cpu_baseline(scene, nsteps):
    """Reference MPMSolver::step (oracle/_ref, built from the reference's own
    sources) on a quarter of the bed, one thread; falls back to the oracle
    restatement when that build is absent."""
    from oracle.pyoracle import OracleSim, Params, RefSim, ref_available

    x = scene.particles["x"]
    lo = x[:, 0].min()
    sel = x[:, 0] < lo + 0.25 * (x[:, 0].max() - lo)
    P = {k: np.asarray(v[sel], np.float64) for k, v in scene.particles.items()}
    prm = Params(E=scene.E, nu=scene.nu, rho=scene.rho, alpha=scene.alpha, mu_floor=scene.mu_floor,
                 gravity=scene.gravity)
    note = None
    if ref_available():
        sim, kind = RefSim(scene.dx, params=prm), "reference"
    else:
        sim, kind = OracleSim(scene.dx, "f32", params=prm), "port"
        note = ("oracle/_ref (the reference's own sources, built where the reference tree is mounted) is absent "
                "on this host: the oracle's FP32 restatement, pinned bit-exact to it, is timed instead")
    sim.set_particles(P)
    sim.step(scene.dt, 1)  # warm-up: allocates the sparse blocks
    t0 = time.perf_counter()
    sim.step(scene.dt, nsteps)
    t = time.perf_counter() - t0
    n = int(P["x"].shape[0])
    out = {"value": n * nsteps / t, "unit": "particle-updates/s", "cores": 1, "kind": kind,
           "sample": f"{n} particles (the quarter of the bed with the lowest x), {nsteps} substeps of "
                     f"MPMSolver::step after 1 warm-up, dt {scene.dt}, single thread",
           "seconds": t, "host_cpu": host_cpu_model()}
    if note:
        out["note"] = note
    return out


def host_cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model or 'unknown'} ({os.cpu_count()} logical CPUs visible)"


def measured_traffic(workload, mode, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary:
    rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch, one run each),
    each scaled by what the same counter reads for a device copy of known
    size.  None when that kernel has not been counted on this workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get("kernels", {}).get(f"{workload}/{mode}/{kernel}")
    if not e:
        return None
    b = e["fetch_kib"] * 1024 * t["fetch_scale"] + e["write_kib"] * 1024 * t["write_scale"]
    return {"bytes": b, "source": TRAFFIC_FILE[len(ROOT) + 1:]}


def _scene_desc(s):
    from basements_amd import M0, M1, SceneDesc

    return SceneDesc(dx=s.dx, origin=s.origin, dims=s.dims, E=s.E, nu=s.nu, rho=s.rho, alpha=s.alpha,
                     mu_floor=s.mu_floor, gravity=s.gravity, walls=s.walls,
                     coupler=M1 if s.tool["method"] == 1 else M0)


def _percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else None


def run_serve(seed, n_requests, rates, cap_p, cap_b, serial=True):
    """Served soil-tool rollouts: a Poisson stream of requests (particle
    counts log-uniform in [5e4, 2e6], tool depth U(0.02, 0.1) m, speed
    U(0.05, 0.5) m/s, substeps uniform in {200..2000}) through the request
    loop at each arrival rate; per rate the p50 / p99 latency (arrival to
    result in host memory) and the aggregate particle-updates/s over the
    makespan (first arrival to last result).  Against the same requests run
    one at a time on single-scene handles (upload, coupled substeps,
    download)."""
    from basements_amd import MODE_FAST, M0, M1, REQ_DONE, Server, Solver
    from basements_amd.scenes import poisson_arrivals, serve_requests, tool_body

    reqs = serve_requests(seed, n_requests)
    dt = 1e-4
    total_pu = float(sum(s.n * k for s, k in reqs))
    cap_s = 64

    def run(rate, overlap=True):
        S = Server(max_particles=cap_p, max_blocks=cap_b, max_scenes=cap_s, dt=dt, overlap=overlap)
        # warm-up: two small requests through the same loop
        w = [S.submit(_scene_desc(s), s.particles, 16, body=tool_body(s.tool)) for s, _ in reqs[:2]]
        while any(S.info(r).state != REQ_DONE for r in w):
            S.poll()
        arr = poisson_arrivals(rate, len(reqs), seed)
        t0 = S.clock() + 0.01
        ids = [S.submit(_scene_desc(s), s.particles, k, body=tool_body(s.tool), arrival=t0 + float(t))
               for (s, k), t in zip(reqs, arr)]
        left = set(ids)
        most = 0
        reported = {}  # id -> server clock when poll handed it back (its output written)
        while left:
            for rid in S.poll():
                reported[rid] = S.clock()
                left.discard(rid)
            most = max(most, S.load()[1])
        infos = [S.info(r) for r in ids]
        bad = [(r, i.state, i.status, i.substeps_done) for r, i in zip(ids, infos) if i.state != REQ_DONE]
        if bad:
            raise RuntimeError(f"requests (id, state, status, substeps done) {bad} did not complete")
        lat = [reported[r] - i.arrival for r, i in zip(ids, infos)]
        makespan = max(reported.values()) - min(i.arrival for i in infos)
        st = S.stats()
        S.close()
        return {"rate": rate, "overlap": overlap, "p50_ms": _percentile(lat, 50) * 1e3,
                "p99_ms": _percentile(lat, 99) * 1e3, "max_ms": max(lat) * 1e3, "makespan_s": makespan,
                "value": total_pu / makespan, "most_running": most, "server_time": st}

    per_rate = [run(float(r)) for r in rates]
    top = max(per_rate, key=lambda r: r["rate"])
    # the top rate again with every poll waiting for its chunk and synchronous staging
    sync = run(top["rate"], overlap=False)

    # the same requests one at a time on the plain path
    t_serial = 0.0
    if serial:
        for s, k in reqs:
            S = Solver(s.dx, s.origin, s.dims, s.n, MODE_FAST)
            S.set_material(s.E, s.nu, s.rho)
            S.set_plastic_alpha(s.alpha)
            S.set_floor_friction(s.mu_floor)
            S.set_gravity(s.gravity)
            S.set_walls(s.walls)
            S.synchronize()
            t0 = time.perf_counter()
            S.upload(s.particles)
            S.set_body(tool_body(s.tool))
            S.step_coupled(M1 if s.tool["method"] == 1 else M0, dt, k)
            S.download()
            t_serial += time.perf_counter() - t0
            S.close()
    serial_v = total_pu / t_serial if t_serial > 0 else None
    return {"reqs": reqs, "total_pu": total_pu, "per_rate": per_rate, "top": top, "sync": sync,
            "serial_value": serial_v,
            "serial_s": t_serial if serial_v else None,
            "config": {"workload": "serve", "requests": len(reqs), "particles": int(sum(s.n for s, _ in reqs)),
                       "particle_updates": total_pu, "particles_min": min(s.n for s, _ in reqs),
                       "particles_max": max(s.n for s, _ in reqs), "substeps_min": min(k for _, k in reqs),
                       "substeps_max": max(k for _, k in reqs), "dt": dt, "server_particles": cap_p,
                       "server_blocks": cap_b, "server_scenes": cap_s}}


def bench_serve(a):
    r = run_serve(a.seed, a.requests, [float(x) for x in a.rates.split(",")], a.serve_particles, a.serve_blocks,
                  bool(a.serial))
    top, serial = r["top"], r["serial_value"]
    out = {"metric": "particle-updates/s (served requests, aggregate)", "value": top["value"],
           "unit": "particle-updates/s", "n_gpus": 1, "steps": len(r["reqs"]), "warmup": 2,
           "ms_per_step": top["makespan_s"] / len(r["reqs"]) * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded soil-tool requests, Poisson arrivals, no external data)",
           "config": r["config"], "rates": r["per_rate"], "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
           "serial_value": serial, "serial_s": r["serial_s"],
           "speedup_vs_serial": top["value"] / serial if serial else None, "synchronous": _sync_summary(r)}
    print(json.dumps(out))
    return 0


def bench_coupler_loops(a):
    """The research harness's grid-only coupler loops (box drop and rover
    wheel, M0, M1, M4.5a, M4.5b and M4.5; 720 steps each: body gravity,
    coupler, reaction, integration, grid settling, all on the device), timed
    without the per-step logging reads.  One JSON line: steps/s over the ten
    loops."""
    from basements_amd import M0, M1, M45, M45A, M45B, harness

    runs = []
    total_steps, total_s = 0, 0.0
    for scen in ("box_drop", "rover_wheel"):
        for name, m in (("M0", M0), ("M1", M1), ("M4.5a", M45A), ("M4.5b", M45B), ("M4.5", M45)):
            sc = harness.SCENARIOS[scen]()
            harness.run(sc, m, steps=min(a.warmup, sc.num_steps()), log=False)  # warm-up
            _, secs = harness.run(sc, m, log=False)
            n = sc.num_steps()
            runs.append({"run": f"{scen}_{name}", "steps": n, "us_per_step": secs / n * 1e6})
            total_steps += n
            total_s += secs
    out = {"metric": "coupler-loop steps/s (grid-only couplers + rigid body)", "value": total_steps / total_s,
           "unit": "steps/s", "n_gpus": 1, "steps": total_steps, "warmup": a.warmup,
           "ms_per_step": total_s / total_steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "the reference's scenario constants (no external data)",
           "config": {"workload": "coupler_loops", "dt": 1.0 / 240.0}, "runs": runs}
    print(json.dumps(out))
    return 0


def bench_adjoint(a):
    """Differentiable tine rollout: 32 x 4 x 32 cells at dx = 1/32 (32,768
    particles), 400 substeps, checkpoint every 16; loss = mean squared final
    displacement against a seeded target field; gradient w.r.t. tool velocity
    (3), friction angle and Young's modulus.  value = particle-substeps per
    second through forward + adjoint in FP32.  The check: the FP64 adjoint
    against FP64 central differences (step 1e-4 relative) and the FP32
    adjoint against the FP64 one."""
    import dataclasses

    from basements_amd import Rollout
    from basements_amd.scenes import tine_scene

    sc = tine_scene(seed=a.seed)
    n, T, dt = sc["particles"]["x"].shape[0], sc["substeps"], sc["dt"]

    def rollout(prec, prm=None):
        R = Rollout(sc["dx"], sc["origin"], sc["dims"], n, precision=prec, checkpoint_every=16)
        R.set_params(prm or sc["params"])
        R.upload(sc["particles"], sc["target"])
        return R

    if not a.single_rollout:
        b = bench_adjoint_rollouts(a, sc, None, None, None)
        print(json.dumps({"metric": f"particle-substeps/s (forward + adjoint, FP32, checkpoint every 16, "
                                    f"{b['rollouts']} rollouts per handle)", "value": b["value"],
                          "unit": "particle-substeps/s", "n_gpus": 1, "steps": max(1, a.steps // 25), "warmup": 1,
                          "ms_per_step": (b["forward_ms"] + b["backward_ms"]) / T, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                          "data": "synthetic (seeded soil bed and target field, no external data)",
                          "config": {"workload": "adjoint", "particles": n, "substeps": T, "dt": dt,
                                     "checkpoint_every": 16, "grid_dims": list(sc["dims"])},
                          "batched_rollouts": b}))
        return 0
    R32 = rollout(32)
    R32.forward(dt, T)
    R32.backward()  # warm-up
    t_f = t_b = 0.0
    reps = max(1, a.steps // 25)
    for _ in range(reps):
        t0 = time.perf_counter()
        loss32 = R32.forward(dt, T)
        t1 = time.perf_counter()
        g32 = R32.backward()
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_b += t2 - t1
    t_f /= reps
    t_b /= reps
    check = None
    if a.fd_check:
        R64 = rollout(64)
        loss64 = R64.forward(dt, T)
        g64 = R64.backward()
        p0 = sc["params"]

        def fd(name, idx=None, rel=1e-4):
            if name == "tool_velocity":
                h = rel * float(np.linalg.norm(p0.tool_velocity))
                vp, vm = list(p0.tool_velocity), list(p0.tool_velocity)
                vp[idx] += h
                vm[idx] -= h
                pp = dataclasses.replace(p0, tool_velocity=tuple(vp))
                pm = dataclasses.replace(p0, tool_velocity=tuple(vm))
            else:
                v = getattr(p0, name)
                h = rel * abs(v)
                pp, pm = dataclasses.replace(p0, **{name: v + h}), dataclasses.replace(p0, **{name: v - h})
            return (rollout(64, pp).forward(dt, T) - rollout(64, pm).forward(dt, T)) / (2 * h)

        fdv = [fd("tool_velocity", k) for k in range(3)]
        fde, fdp = fd("young_modulus"), fd("friction_deg")
        adj = [*g64["tool_velocity"], g64["young_modulus"], g64["friction_deg"]]
        fds = [*fdv, fde, fdp]
        vscale = max(abs(v) for v in fdv)
        rel = [abs(x - y) / vscale for x, y in zip(adj[:3], fds[:3])] + \
              [abs(x - y) / abs(y) for x, y in zip(adj[3:], fds[3:])]
        g32v = [*g32["tool_velocity"], g32["young_modulus"], g32["friction_deg"]]
        rel32 = [abs(x - y) / vscale for x, y in zip(g32v[:3], adj[:3])] + \
                [abs(x - y) / abs(y) for x, y in zip(g32v[3:], adj[3:])]
        check = {"params": ["tool_vx", "tool_vy", "tool_vz", "young_modulus", "friction_deg"],
                 "adjoint_fp64": adj, "central_fd_fp64": fds, "rel_err_fp64_vs_fd": rel,
                 "adjoint_fp32": g32v, "rel_err_fp32_vs_fp64": rel32, "loss_fp64": loss64, "loss_fp32": loss32,
                 "fd_step": "1e-4 relative (tool speed for velocity components)"}
    value = n * T / (t_f + t_b)
    batched = bench_adjoint_rollouts(a, sc, loss32, g32, value) if a.rollouts > 1 else None
    # HBM model: the forward substep's 200 B per particle (state read and
    # written once), the backward twice that (the recomputed forward state plus
    # the adjoint state read and written); the bound at this size is occupancy
    waves = -(-n // 64)
    algo = ADJ_BYTES_PER_PARTICLE_SUBSTEP * n * T
    roof = {"bound": "latency", "achieved": algo / (t_f + t_b) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": algo / (t_f + t_b) / 1e9 / HBM_PEAK_GBS, "traffic": None,
            "algo_bytes_per_particle_substep": ADJ_BYTES_PER_PARTICLE_SUBSTEP, "waves_per_kernel": waves,
            "simds": 1024, "waves_per_simd": waves / 1024}
    out = {"metric": "particle-substeps/s (forward + adjoint, FP32, checkpoint every 16)", "value": value,
           "unit": "particle-substeps/s", "n_gpus": 1, "steps": reps, "warmup": 1,
           "ms_per_step": (t_f + t_b) / T * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32", "data": "synthetic (seeded soil bed and target field, no external data)",
           "config": {"workload": "adjoint", "particles": n, "substeps": T, "dt": dt, "checkpoint_every": 16,
                      "grid_dims": list(sc["dims"]), "tool_velocity": list(sc["params"].tool_velocity)},
           "forward_ms": t_f * 1e3, "backward_ms": t_b * 1e3, "roofline": roof, "check": check,
           "batched_rollouts": batched}
    print(json.dumps(out))
    return 0


def bench_adjoint_rollouts(a, sc, loss_single, g_single, single_rate):
    """A parameter sweep as one handle of R rollouts (bm_diff_desc.rollouts):
    rollout r has Young's modulus x (1 + r/R) and friction angle + r/2 deg,
    rollout 0 the scene's own parameters.  Rate = particle-substeps of all
    rollouts per second through forward + adjoint in FP32; check = rollout 0
    against the single-rollout handle's loss and gradient."""
    import dataclasses

    from basements_amd import Rollout

    nr = a.rollouts
    n, T, dt = sc["particles"]["x"].shape[0], sc["substeps"], sc["dt"]
    p0 = sc["params"]
    B = Rollout(sc["dx"], sc["origin"], sc["dims"], n, precision=32, checkpoint_every=16, rollouts=nr)
    for r in range(nr):
        B.set_params(dataclasses.replace(p0, young_modulus=p0.young_modulus * (1 + r / nr),
                                         friction_deg=p0.friction_deg + 0.5 * r), rollout=r)
    B.upload(sc["particles"], sc["target"])
    B.forward_all(dt, T)
    B.backward_all()  # warm-up
    reps = max(1, a.steps // 25)
    t_f = t_b = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        losses = B.forward_all(dt, T)
        t1 = time.perf_counter()
        grads = B.backward_all()
        t2 = time.perf_counter()
        t_f += t1 - t0
        t_b += t2 - t1
    t_f /= reps
    t_b /= reps
    rate = nr * n * T / (t_f + t_b)
    out = {"rollouts": nr, "value": rate, "unit": "particle-substeps/s", "forward_ms": t_f * 1e3,
           "backward_ms": t_b * 1e3, "vs_single_rollout": rate / single_rate if single_rate else None,
           "sweep": "young_modulus x (1 + r/R), friction_deg + r/2", "losses": [float(x) for x in losses]}
    if g_single is None:
        return out
    g0 = grads[0]
    vscale = float(np.abs(g_single["tool_velocity"]).max())
    return {**out,
            "rollout0_rel_loss_vs_single": abs(losses[0] - loss_single) / abs(loss_single),
            "rollout0_rel_grad_vs_single": {
                "young_modulus": abs(g0["young_modulus"] - g_single["young_modulus"]) / abs(g_single["young_modulus"]),
                "friction_deg": abs(g0["friction_deg"] - g_single["friction_deg"]) / abs(g_single["friction_deg"]),
                "tool_velocity": float(np.abs(g0["tool_velocity"] - g_single["tool_velocity"]).max()) / vscale}}


def run_batch(seed, n_scenes, steps, warmup, graph=True, tuning="", serial=True):
    """Batched uneven scenes: aggregate particle-updates/s of all scenes in one
    batch (autotuned, graph-replayed) against running the same scenes one at a
    time, each on its own single-scene handle, for the same substeps."""
    from basements_amd import M0, M1, MODE_FAST, Batch, Solver
    from basements_amd.scenes import batch_scenes, tool_body

    scenes = batch_scenes(seed, n_scenes)
    dt = 1e-4
    n_total = sum(s.n for s in scenes)
    blocks = sum((-(-s.dims[0] // 4)) * (-(-s.dims[1] // 4)) * (-(-s.dims[2] // 4)) for s in scenes)

    B = Batch(max_particles=n_total + 64, max_blocks=blocks, max_scenes=len(scenes))
    ids = [B.add_scene(_scene_desc(s), s.particles, tool_body(s.tool)) for s in scenes]
    if tuning:
        unit, sort_every, groups = (int(t) for t in tuning.split(","))
        B.set_tuning(unit, sort_every, groups)
    else:
        unit, sort_every, groups = B.autotune(dt, substeps=64)
    B.set_graph(bool(graph))
    # warm-up long enough to capture the replayed substeps (two re-binning cycles)
    warmup = max(warmup, 3 * sort_every)
    B.step(dt, warmup)
    B.synchronize()
    t0 = time.perf_counter()
    B.step(dt, steps)
    B.synchronize()
    t_batch = time.perf_counter() - t0
    del B

    # the same scenes one at a time on the plain path
    t_serial = 0.0
    for s in scenes if serial else []:
        S = Solver(s.dx, s.origin, s.dims, s.n, MODE_FAST)
        S.set_material(s.E, s.nu, s.rho)
        S.set_plastic_alpha(s.alpha)
        S.set_floor_friction(s.mu_floor)
        S.set_gravity(s.gravity)
        S.upload(s.particles)
        S.set_body(tool_body(s.tool))
        method = M1 if s.tool["method"] == 1 else M0
        S.step_coupled(method, dt, warmup)
        S.synchronize()
        t0 = time.perf_counter()
        S.step_coupled(method, dt, steps)
        S.synchronize()
        t_serial += time.perf_counter() - t0
        S.close()
    value = n_total * steps / t_batch
    serial_v = n_total * steps / t_serial if t_serial > 0 else None
    return {"value": value, "steps": steps, "warmup": warmup, "ms_per_step": t_batch / steps * 1e3,
            "serial_value": serial_v, "serial_ms_per_step": t_serial / steps * 1e3 if serial_v else None,
            "speedup_vs_serial": value / serial_v if serial_v else None,
            "config": {"workload": "batch", "scenes": len(scenes), "particles": n_total, "grid_blocks": blocks,
                       "particles_min": min(s.n for s in scenes), "particles_max": max(s.n for s in scenes),
                       "grids": sorted({s.dims[0] for s in scenes}), "dt": dt, "graph": bool(graph),
                       "tuned_unit_particles": unit, "tuned_sort_every": sort_every,
                       "tuned_deposit_groups": groups}}


def bench_batch(a):
    r = run_batch(a.seed, a.scenes, a.steps, a.warmup, bool(a.graph), a.tuning, bool(a.serial))
    out = {"metric": "particle-updates/s (batched scenes, aggregate)", "value": r["value"],
           "unit": "particle-updates/s", "n_gpus": 1, "steps": r["steps"], "warmup": r["warmup"],
           "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32", "data": "synthetic (seeded soil beds with dragged tools, no external data)",
           "config": r["config"], "serial_value": r["serial_value"], "serial_ms_per_step": r["serial_ms_per_step"],
           "speedup_vs_serial": r["speedup_vs_serial"]}
    print(json.dumps(out))
    return 0


def main():
    a = _args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build has no multi-GPU path)",
              file=sys.stderr)
        return 2
    if a.copy_calibration:
        copy_calibration()
        return 0
    if a.config == "batch":
        return bench_batch(a)
    if a.config == "serve":
        return bench_serve(a)
    if a.config == "adjoint":
        return bench_adjoint(a)
    if a.config == "coupler_loops":
        return bench_coupler_loops(a)

    from basements_amd import M0, M1, MODE_FAST, MODE_REFERENCE, Solver
    from basements_amd.scenes import box_drop, wheel_bed

    scene = wheel_bed(a.seed) if a.config == "wheel_bed" else box_drop(a.seed)
    mode = MODE_FAST if a.mode == "fast" else MODE_REFERENCE
    S = Solver(scene.dx, scene.origin, scene.dims, scene.n, mode)
    S.set_material(scene.E, scene.nu, scene.rho)
    S.set_plastic_alpha(scene.alpha)
    S.set_floor_friction(scene.mu_floor)
    S.set_gravity(scene.gravity)
    S.upload(scene.particles)
    S.synchronize()
    if a.tuning and mode == MODE_FAST:
        unit, sort_every = (int(t) for t in a.tuning.split(",")[:2])
        S.set_tuning(unit, sort_every, -1)

    dt = scene.dt
    tool = scene.tool
    method = M1 if tool["method"] == 1 else M0
    # One step = p2g, grid update, coupler with the reaction into the device
    # body and its integration, g2p (bm_mpm_step_coupled): no host round trip.
    S.set_body(make_body(tool))

    # Fast mode re-bins the particles every REBIN_PERIOD substeps.  The
    # warm-up is rounded up to whole periods, so the timed window starts with
    # a re-binning and holds ceil(K / period) of them -- at least its share
    # of the steady state, whatever K is.
    fast = mode == MODE_FAST
    period = (int(a.tuning.split(",")[1]) if a.tuning and mode == MODE_FAST else REBIN_PERIOD) if fast else 0
    warm = -(-a.warmup // period) * period if fast else a.warmup
    S.step_coupled(method, dt, warm)
    S.synchronize()
    # Timed region: the G2P kernel -- the dominant one -- carries HIP events
    # recorded by its own dispatches (fast mode), nothing is recorded between
    # launches; the other kernels are timed in the phase pass below (each
    # recorded dispatch lengthens the substep by about 5 us).
    S.set_timing(True, g2p_only=fast)
    S.read_timing(reset=True)
    t0 = time.perf_counter()
    S.step_coupled(method, dt, a.steps)
    S.synchronize()
    t = time.perf_counter() - t0
    tk = S.read_timing(reset=True)
    # Phase breakdown: a further, untimed pass with event pairs around every
    # launch (they add their own gaps, so this pass is not the throughput).
    n_phase = min(a.steps, 20)
    if fast:
        S.set_timing(True)
        S.step_coupled(method, dt, n_phase)
        S.synchronize()
        tm = S.read_timing(reset=True)
    else:
        tm, n_phase = tk, a.steps
    S.set_timing(False)

    # active nodes of the last substep (mass > 0)
    G = S.download_grid()
    n_nodes = int((G["mass_q"] != 0).sum())
    del G
    st = S.stats()
    n = scene.n
    value = n * a.steps / t

    per_kernel = {}
    node_bytes = 80 if mode == MODE_REFERENCE else 56
    for k, (bp, bn) in (BYTES_REF if mode == MODE_REFERENCE else BYTES_FAST).items():
        ms, calls = tk.get(k, (0.0, 0))
        src = "timed region"
        if not calls:
            ms, calls = tm.get(k, (0.0, 0))
            src = "phase pass"
        if calls:
            avg_s = ms / calls * 1e-3
            algo = bp * n + bn * n_nodes
            per_kernel[k] = {"avg_ms": ms / calls, "launches": calls, "algo_bytes": algo,
                             "achieved_GBs": algo / avg_s / 1e9, "timed_in": src}
    for k, v in per_kernel.items():
        v["kernel"] = KERNEL_OF.get((a.mode, k), k)
    # The dominant kernel: the longest in the phase pass, where every kernel
    # is timed the same way (the timed region records only the G2P's
    # dispatches, so its figures and the phase pass's are not comparable).
    dom = max(per_kernel, key=lambda k: (tm.get(k, (0.0, 0))[0] / max(tm.get(k, (0.0, 0))[1], 1), per_kernel[k]["avg_ms"]))
    ach = per_kernel[dom]["achieved_GBs"]
    sub_bytes = 200 * n + node_bytes * n_nodes
    gpu_ms = sum(ms for ms, _ in tm.values()) / n_phase
    traffic = measured_traffic(scene.name, a.mode, per_kernel[dom]["kernel"])

    out = {
        "metric": "particle-updates/s",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered soil bed, no external data)",
        "config": {"workload": scene.name, "particles": n, "grid": list(scene.dims), "dx": scene.dx,
                   "dt": dt, "coupler": "M1" if method == M1 else "M0", "mode": a.mode,
                   "active_nodes": n_nodes, "active_blocks": st["active_blocks"],
                   "rebin_period": period or None, "rebins_in_window": -(-a.steps // period) if fast else None,
                   "warmup_rounded_to_period": warm},
        "roofline": {"bound": "hbm", "kernel": per_kernel[dom]["kernel"], "phase": dom, "achieved": ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "algo_bytes_per_launch": per_kernel[dom]["algo_bytes"],
                     "avg_ms": per_kernel[dom]["avg_ms"]},
        # the whole substep at the timed window's wall time per substep; the
        # phase pass's summed event pairs (gpu_ms) carry their own gaps
        "substep_roofline": {"algo_bytes": sub_bytes, "bytes_per_node": node_bytes, "ms": t / a.steps * 1e3,
                             "achieved_GBs": sub_bytes / (t / a.steps) / 1e9,
                             "frac": sub_bytes / (t / a.steps) / 1e9 / HBM_PEAK_GBS,
                             "phase_pass_gpu_ms": gpu_ms},
        "kernels": per_kernel,
        "phase_ms_per_step": {k: ms / n_phase for k, (ms, calls) in tm.items() if calls},
        "phase_pass_substeps": n_phase,
        "fp64_svd_fallbacks": st["fp64_fallbacks"],
    }
    if a.cpu_baseline:
        cb = cpu_baseline(scene, a.cpu_steps)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "host_cpu", "note")
                               if k in cb}
    S.close()
    del S
    if a.batched and a.config == "wheel_bed":
        r = run_batch(a.seed, 64, BATCH_STEPS, 10)
        out["batched"] = {"metric": "particle-updates/s (64 batched scenes, aggregate)", "value": r["value"],
                          "serial_value": r["serial_value"], "speedup_vs_serial": r["speedup_vs_serial"],
                          "steps": r["steps"], "ms_per_step": r["ms_per_step"],
                          "serial_ms_per_step": r["serial_ms_per_step"], "config": r["config"]}
    if a.served and a.config == "wheel_bed":
        r = run_serve(a.seed, 48, [SERVE_RATE], a.serve_particles, a.serve_blocks)
        top = r["top"]
        out["served"] = {"metric": "particle-updates/s (served requests, aggregate over the makespan)",
                         "value": top["value"], "rate_per_s": top["rate"], "p50_ms": top["p50_ms"],
                         "p99_ms": top["p99_ms"], "makespan_s": top["makespan_s"],
                         "serial_value": r["serial_value"],
                         "speedup_vs_serial": top["value"] / r["serial_value"] if r["serial_value"] else None,
                         "most_running": top["most_running"], "server_time": top["server_time"],
                         "synchronous": _sync_summary(r), "config": r["config"]}
        if "batched" in out:
            # if the device ran at the 64-scene batch's rate from the first
            # arrival to the last result, with nothing exposed
            floor_s = r["total_pu"] / out["batched"]["value"]
            out["served"]["device_bound_makespan_s"] = floor_s
            out["served"]["speedup_ceiling"] = r["serial_s"] / floor_s if r["serial_s"] else None
    print(json.dumps(out))
    return 0


def _sync_summary(r):
    """The served run repeated with overlap off (each poll waits for its
    chunk; staging on the host path between chunks)."""
    y, top = r["sync"], r["top"]
    return {"value": y["value"], "makespan_s": y["makespan_s"], "p50_ms": y["p50_ms"], "p99_ms": y["p99_ms"],
            "server_time": y["server_time"], "overlap_speedup": y["makespan_s"] / top["makespan_s"]}


if __name__ == "__main__":
    sys.exit(main())
