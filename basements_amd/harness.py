"""Grid-only coupler loops of the reference's research harness on the device.

The reference's ``bench_compare_all`` drives its rigid-body couplers on a
seeded sand grid with no particles: every step applies gravity to the body,
couples it (M0 or M1), adds the reaction, integrates the body, then lets the
grid settle (v_y += -g dt, clamped at 0) and logs one time-series row
(``research/mpm_rigid_coupling/bench_compare_all.cpp:72-134`` box drop,
``:614-689`` rover wheel; the row format is ``FrameExporter::log_step``,
``frame_exporter.h:44-57``).  Here the coupling, reaction, integration and
settling run on the GPU (``bm_mpm_couple_body``, ``bm_mpm_grid_settle``);
the host reads the body and the grid back only to log.

Scenario constants follow ``scenarios/scenario_box_drop.h:14-78`` and
``scenarios/scenario_rover_wheel.h:19-176``.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

COLUMNS = ("t", "box_x", "box_y", "box_z", "box_vy", "p_body_y", "p_grid_y", "p_sum_y", "KE", "nodes_inside")


@dataclass
class GridOnlyScenario:
    """A seeded sand patch (nodes i in [-ei, ei], j in [0, layers), k in [-ek,
    ek], each with mass ``node_mass`` at rest) and one dynamic body."""

    name: str
    dx: float
    extent_i: int
    extent_k: int
    layers_y: int
    node_mass: float
    shape: int            # SHAPE_BOX or SHAPE_CYLINDER
    half_extents: tuple
    body_mass: float
    position: tuple
    orientation: tuple
    linear_velocity: tuple
    angular_velocity: tuple
    friction: float = 0.5
    dt: float = 1.0 / 240.0
    total_time: float = 3.0
    gravity_g: float = 9.81
    frame_every_n: int = 5  # a grid snapshot every n steps (scenario_box_drop.h:35)
    settle_time: float | None = None  # foot step: rebound samples start here (scenario_foot_step.h:49)
    # the node box the handle holds: covers the patch, the body's reach and
    # block alignment with the reference's 4^3 blocks (origin a multiple of 4)
    origin: tuple = (-16, -16, -16)
    dims: tuple = (32, 32, 32)

    def num_steps(self) -> int:
        # scenario_box_drop.h:37-39 (same in the rover wheel)
        return int(np.float32(self.total_time) / np.float32(self.dt) + np.float32(0.5))


def box_drop() -> GridOnlyScenario:
    """ScenarioBoxDrop (scenario_box_drop.h:14-78): a 1 kg box, half extents
    0.2 m, dropped from y = 0.55 m on a 7 x 3 x 7 node patch at dx = 0.1 m.
    make_box sets inertia_tensor directly, so the inverse inertia stays the
    identity (rigid_body.h:122)."""
    from .mpm import SHAPE_BOX

    return GridOnlyScenario(name="box_drop", dx=0.1, extent_i=3, extent_k=3, layers_y=3, node_mass=0.01,
                            shape=SHAPE_BOX, half_extents=(0.2, 0.2, 0.2), body_mass=1.0, position=(0.0, 0.55, 0.0),
                            orientation=(1.0, 0.0, 0.0, 0.0), linear_velocity=(0.0, 0.0, 0.0),
                            angular_velocity=(0.0, 0.0, 0.0))


def rover_wheel() -> GridOnlyScenario:
    """ScenarioRoverWheel in ROLL mode (scenario_rover_wheel.h:19-176): a 1 kg
    cylinder R = 0.15 m, half length 0.05 m, its axis turned onto world Z
    (Quaternion::from_axis_angle((1,0,0), 1.5707963f) in FP32, quaternion.h:36-47),
    spinning at -10 rad/s about Z, on a 13 x 3 x 9 node patch."""
    from .mpm import SHAPE_CYLINDER

    half = np.float32(1.5707963) * np.float32(0.5)
    s, c = np.float32(math.sin(float(half))), np.float32(math.cos(float(half)))
    return GridOnlyScenario(name="rover_wheel", dx=0.1, extent_i=6, extent_k=4, layers_y=3, node_mass=0.01,
                            shape=SHAPE_CYLINDER, half_extents=(0.15, 0.05, 0.15), body_mass=1.0,
                            position=(0.0, 0.55, 0.0), orientation=(float(c), float(s), 0.0, 0.0),
                            linear_velocity=(0.0, 0.0, 0.0), angular_velocity=(0.0, 0.0, -10.0))


def box_drop_dense() -> GridOnlyScenario:
    """The box drop with a 9 kg box (bench_compare_all.cpp:505-509:
    S.box_mass = 9, density-matched to the wheel)."""
    sc = box_drop()
    sc.name, sc.body_mass = "box_drop_dense", 9.0
    return sc


def foot_step() -> GridOnlyScenario:
    """ScenarioFootStep (scenario_foot_step.h): a 1 kg thin box foot, half
    extents (0.10, 0.02, 0.10) m, dropped from y = 0.55 m on the 7 x 3 x 7
    patch, 2 s.  make_foot sets inertia_tensor directly: the inverse inertia
    stays the identity."""
    sc = box_drop()
    sc.name, sc.half_extents, sc.total_time = "foot_step", (0.10, 0.02, 0.10), 2.0
    sc.settle_time = 0.05
    return sc


SCENARIOS = {"box_drop": box_drop, "rover_wheel": rover_wheel, "box_drop_dense": box_drop_dense,
             "foot_step": foot_step}

class ReboundTracker:
    """The foot step's rebound coefficient (ReboundTracker,
    scenario_foot_step.h:91-137), observed on the body's v_y after each step
    from settle_time on: the most negative v_y while descending is
    v_down_peak; once v_y turns non-negative after it, the largest v_y until
    v_y is non-positive again is v_up_peak; e = v_up_peak / |v_down_peak|,
    0 when the body never descended or never rose, capped at 2.  All in FP32
    as the harness keeps it."""

    DESCENT, REBOUND, DONE = 0, 1, 2
    COLUMNS = ("v_down_peak", "v_up_peak", "rebound_coefficient_e")

    def __init__(self):
        self.state = self.DESCENT
        self.v_down = np.float32(0.0)
        self.v_up = np.float32(0.0)

    def observe(self, t: float, v_y: float, settle_time: float) -> None:
        v = np.float32(v_y)
        if np.float32(t) < np.float32(settle_time):
            return
        if self.state == self.DESCENT:
            self.v_down = min(self.v_down, v)
            if v >= 0 and self.v_down < 0:
                self.state = self.REBOUND
        if self.state == self.REBOUND:
            self.v_up = max(self.v_up, v)
            if v <= 0 and self.v_up > 0:
                self.state = self.DONE

    def coefficient(self) -> float:
        if self.v_down >= 0:
            return 0.0
        e = self.v_up / -self.v_down
        return float(min(max(e, np.float32(0.0)), np.float32(2.0)))

    def summary(self) -> dict:
        return {"v_down_peak": float(self.v_down), "v_up_peak": float(self.v_up),
                "rebound_coefficient_e": self.coefficient()}


METRIC_COLUMNS = ("max_abs_total_p_y", "avg_abs_total_p_y", "max_abs_energy_drift", "step_time_avg_us",
                  "max_penetration_m", "avg_penetration_m", "dwell_fraction", "ratio_p_y_vs_M0")
PENETRATION_DWELL_REL = np.float32(0.30)  # kPenetrationDwellRelThreshold, metrics.h


class RunMetrics:
    """The harness's run summary (RunMetrics, metrics.h): peak and mean |total
    p_y|, peak relative energy drift against the first step, peak and mean
    penetration depth and the fraction of steps deeper than 0.3 of the smallest
    half extent; FP32 accumulators as there."""

    def __init__(self):
        self.max_p = np.float32(0)
        self.sum_p = np.float32(0)
        self.max_drift = np.float32(0)
        self.e0 = None
        self.time_us = 0.0
        self.n = 0
        self.max_pen = np.float32(0)
        self.sum_pen = np.float32(0)
        self.dwell = 0

    def observe(self, total_p_y, e_total, pen, min_half, step_us):
        e = np.float32(e_total)
        if self.e0 is None:
            self.e0 = e
        drift = (np.float32(abs(e - self.e0)) / np.float32(abs(self.e0)) if abs(self.e0) > np.float32(1e-9)
                 else np.float32(abs(e - self.e0)))
        ap = np.float32(abs(np.float32(total_p_y)))
        self.max_p = max(self.max_p, ap)
        self.sum_p = np.float32(self.sum_p + ap)
        self.max_drift = max(self.max_drift, drift)
        pen = np.float32(pen)
        self.max_pen = max(self.max_pen, pen)
        self.sum_pen = np.float32(self.sum_pen + pen)
        if pen > np.float32(min_half) * PENETRATION_DWELL_REL:
            self.dwell += 1
        self.time_us += step_us
        self.n += 1

    def summary(self) -> dict:
        n = max(self.n, 1)
        return {"max_abs_total_p_y": float(self.max_p), "avg_abs_total_p_y": float(np.float32(self.sum_p / n)),
                "max_abs_energy_drift": float(self.max_drift), "step_time_avg_us": self.time_us / n,
                "max_penetration_m": float(self.max_pen), "avg_penetration_m": float(np.float32(self.sum_pen / n)),
                "dwell_fraction": self.dwell / n, "ratio_p_y_vs_M0": 1.0}


def _rotation(q) -> np.ndarray:
    """Matrix3::from_quaternion (matrix3.h:432-448) in FP32, row-major."""
    w, x, y, z = (np.float32(v) for v in q)
    two, one = np.float32(2), np.float32(1)
    return np.array([[one - two * (y * y + z * z), two * (x * y - w * z), two * (x * z + w * y)],
                     [two * (x * y + w * z), one - two * (x * x + z * z), two * (y * z - w * x)],
                     [two * (x * z - w * y), two * (y * z + w * x), one - two * (x * x + y * y)]], np.float32)


def penetration(sc: "GridOnlyScenario", position, orientation, nodes_xyz, mass) -> float:
    """measure_penetration (metrics.h): the deepest node (mass > 0) inside the
    body's half-extent box, depth = the smallest distance to a face."""
    if nodes_xyz.shape[0] == 0:
        return 0.0
    R = _rotation(orientation)
    local = (nodes_xyz - np.asarray(position, np.float32)) @ R  # R_w2l (x - c)
    he = np.asarray(sc.half_extents, np.float32)
    d = he - np.abs(local)
    ok = (d >= 0).all(axis=1) & (mass > 0)
    return float(d[ok].min(axis=1).max()) if ok.any() else 0.0


def _seed(sc: GridOnlyScenario):
    """seed_sand / seed_sand_for_wheel as a dense box at the handle's origin."""
    dims = tuple(sc.dims)
    mass = np.zeros(dims, np.float32)
    vel = np.zeros(dims + (3,), np.float32)
    o = sc.origin
    for i in range(-sc.extent_i, sc.extent_i + 1):
        for j in range(sc.layers_y):
            for k in range(-sc.extent_k, sc.extent_k + 1):
                mass[i - o[0], j - o[1], k - o[2]] = np.float32(sc.node_mass)
    return mass, vel, mass > 0


def run(sc: GridOnlyScenario, method: int, steps: int | None = None, log: bool = True, frames: bool = False):
    """The grid-only coupled loop on a fresh handle.  Returns (columns,
    seconds): the time-series rows of FrameExporter::log_step (dict of
    arrays; also "p_grid_abs" = sum |v_y m| and "n_grid" = active nodes, the
    size of the grid momentum sum, for its rounding bound) and the wall time
    of the stepping alone when log is False (the per-step host reads of the
    logging otherwise dominate).  frames: also the snapshots of
    FrameExporter::log_frame every frame_every_n steps, rows["frames"] = list
    of (t, body position, body velocity, nodes) with nodes an (n, 7) array of
    x, y, z (node centres), vx, vy, vz, mass of the nodes with mass > 0.
    A scenario with a settle_time (the foot step) also returns
    rows["rebound"], the ReboundTracker summary of the logged run."""
    from .mpm import BODY_DYNAMIC, MODE_REFERENCE, Body, Solver

    n = sc.num_steps() if steps is None else int(steps)
    S = Solver(sc.dx, sc.origin, sc.dims, 1, MODE_REFERENCE)
    mass, vel, active = _seed(sc)
    S.upload_grid(sc.origin, mass, vel, active)
    m = float(sc.body_mass)
    # RigidBody::apply_force(gravity) each step: a constant applied force;
    # make_dynamic(m) sets inv_mass = 1 / m; damping 0 as make_box / make_wheel
    S.set_body(Body(position=sc.position, orientation=sc.orientation, linear_velocity=sc.linear_velocity,
                    angular_velocity=sc.angular_velocity, inv_mass=1.0 / m, linear_damping=0.0,
                    angular_damping=0.0, integrate=BODY_DYNAMIC,
                    applied_force=(0.0, float(np.float32(-sc.gravity_g) * np.float32(m)), 0.0),
                    half_extents=sc.half_extents, friction=sc.friction, mass=m, shape=sc.shape))
    dvy = float(np.float32(-sc.gravity_g) * np.float32(sc.dt))
    dt = float(np.float32(sc.dt))
    rows = {k: np.zeros(n, np.float64) for k in COLUMNS + ("p_grid_abs", "n_grid")}
    metrics = RunMetrics()
    rebound = ReboundTracker() if sc.settle_time is not None else None
    g32 = np.float32(sc.gravity_g)
    min_half = float(min(np.float32(h) for h in sc.half_extents))
    ii, jj, kk = np.meshgrid(*(np.arange(o, o + d) for o, d in zip(sc.origin, sc.dims)), indexing="ij")
    node_xyz = ((np.stack([ii, jj, kk], -1).astype(np.float32) + np.float32(0.5)) * np.float32(sc.dx))
    # the harness measures penetration with its collider state, whose centre it
    # updates after the integration but whose orientation it leaves at the
    # step's start (bench_compare_all.cpp:126-134, 676-680)
    q_step = tuple(sc.orientation)
    snaps = []
    t0 = time.perf_counter()
    for i in range(n):
        ts = time.perf_counter()
        S.couple_body(method, dt)
        S.grid_settle(dvy)
        if not log:
            continue
        body, (force, torque, inside) = S.get_body()
        G = S.download_grid()
        act = G["active"].astype(bool) & (G["mass"] > 0)
        terms = G["vel"][..., 1][act].astype(np.float64) * G["mass"][act]
        p_grid_y = float(np.sum(terms))
        rows["p_grid_abs"][i] = float(np.abs(terms).sum())
        rows["n_grid"][i] = terms.size
        v = np.asarray(body.linear_velocity, np.float32)
        p_body_y = float(v[1] * np.float32(m))
        rows["t"][i] = float(np.float32(i + 1) * np.float32(sc.dt))
        rows["box_x"][i], rows["box_y"][i], rows["box_z"][i] = body.position
        rows["box_vy"][i] = float(v[1])
        rows["p_body_y"][i] = p_body_y
        rows["p_grid_y"][i] = p_grid_y
        rows["p_sum_y"][i] = p_body_y + p_grid_y
        rows["KE"][i] = float(np.float32(0.5) * np.float32(m) * np.float32(np.dot(v, v)))
        rows["nodes_inside"][i] = inside
        if rebound is not None:  # bench_compare_all.cpp:431, on the body's v_y after the step
            rebound.observe(rows["t"][i], v[1], sc.settle_time)
        # RunMetrics inputs (bench_compare_all.cpp:137-156, 677-690)
        vn = G["vel"][act].astype(np.float64)
        e_grid = float(np.sum(0.5 * G["mass"][act] * np.einsum("ij,ij->i", vn, vn)))
        e_total = float(np.float32(rows["KE"][i]) + np.float32(m) * g32 * np.float32(body.position[1])) + e_grid
        pen = penetration(sc, body.position, q_step, node_xyz[act], G["mass"][act])
        q_step = tuple(body.orientation)
        if frames and i % sc.frame_every_n == 0:  # bench_compare_all.cpp:132
            nodes = np.concatenate([node_xyz[act], G["vel"][act], G["mass"][act][:, None]], 1).astype(np.float32)
            snaps.append((rows["t"][i], tuple(body.position), tuple(body.linear_velocity), nodes))
        metrics.observe(rows["p_sum_y"][i], e_total, pen, min_half, (time.perf_counter() - ts) * 1e6)
    S.synchronize()
    secs = time.perf_counter() - t0
    S.close()
    if log:
        rows["metrics"] = metrics.summary()
        if rebound is not None:
            rows["rebound"] = rebound.summary()
    if frames:
        rows["frames"] = snaps
    return rows, secs


def write_frames_csv(path: str, frames: list) -> None:
    """FrameExporter::log_frame's format (frame_exporter.h:59-94): per frame a
    body row and one row per node with mass > 0 (here in block-linear order;
    the reference's follows its hash map)."""
    with open(path, "w") as f:
        f.write("t,kind,id,x,y,z,vx,vy,vz,mass\n")
        for t, pos, vel, nodes in frames:
            f.write("%.6f,B,0,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e\n" % ((t,) + tuple(pos) + tuple(vel) + (0.0,)))
            for k, r in enumerate(nodes):
                f.write("%.6f,N,%d,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e\n" % ((t, k) + tuple(float(v) for v in r)))


TINE_COLUMNS = ("t", "tine_x", "reaction_fx", "reaction_fy", "reaction_fz")


def method_id(name: str) -> int:
    """The coupler of a harness method name.  The harness's M2 (ASFLIP), M3
    (MLS) and M4 (Lagrange) are skeletons that delegate to M1
    (couplers/m2_asflip.h, m3_mls_mpm.h, m4_lagrange.h: apply returns
    MPMRigidCouplerPIC::apply), so their committed runs are M1's."""
    from .mpm import M0, M1, M45, M45A, M45B

    return {"M0": M0, "M1": M1, "M2": M1, "M3": M1, "M4": M1, "M4.5a": M45A, "M4.5b": M45B, "M4.5": M45}[name]


def run_tine_drag(method: int, steps: int | None = None):
    """ScenarioTineDrag (scenario_tine_drag.h, run_tine_drag in
    bench_compare_all.cpp:215-292): a 10 x 100 x 80 mm tine dragged at
    0.2 m/s through a 13 x 4 x 7 node soil bed (dx 5 cm, 0.1875 kg per node),
    6 cm deep, kinematically; per step the tine moves, the coupler acts and
    the grid settles.  Returns (rows of every 4th step: t, tine_x and the
    reaction, the summary: mean draft -F_x and vertical F_y after 0.5 s,
    specific draught per metre of tine width, samples).  On the device the
    tine is a kinematic body, which moves after its coupling: it starts one
    step ahead, so it couples at the reference's positions, accumulated the
    same way in FP32."""
    from .mpm import BODY_KINEMATIC, MODE_REFERENCE, SHAPE_BOX, Body, Solver

    f32 = np.float32
    dx, dt, g, v = f32(0.05), f32(1.0) / f32(240.0), f32(9.81), f32(0.2)
    he = (0.005, 0.05, 0.04)
    n = int(f32(2.0) / dt + f32(0.5)) if steps is None else int(steps)
    origin, dims = (-8, -8, -8), (32, 16, 16)
    mass = np.zeros(dims, np.float32)
    for i in range(-2, 11):
        for j in range(4):
            for k in range(-3, 4):
                mass[i - origin[0], j - origin[1], k - origin[2]] = f32(0.1875)
    S = Solver(float(dx), origin, dims, 1, MODE_REFERENCE)
    S.upload_grid(origin, mass, np.zeros(dims + (3,), np.float32), mass > 0)
    y0 = (f32(4 - 1) + f32(0.5)) * dx - f32(0.06) + f32(he[1])  # tine_center_y_init
    x1 = f32(-0.05) + v * dt  # tine.position.x += v dt comes first in each reference step
    S.set_body(Body(position=(float(x1), float(y0), 0.0), orientation=(1.0, 0.0, 0.0, 0.0),
                    linear_velocity=(float(v), 0.0, 0.0), angular_velocity=(0.0, 0.0, 0.0), inv_mass=2.0,
                    linear_damping=0.0, angular_damping=0.0, integrate=BODY_KINEMATIC, half_extents=he,
                    friction=0.55, mass=0.5, shape=SHAPE_BOX))
    dvy = float(-g * dt)
    rows, fx_sum, fy_sum, samples = [], 0.0, 0.0, 0
    for i in range(n):
        t = float(f32(i + 1) * dt)
        x_i = S.get_body()[0].position[0] if i % 4 == 0 else None  # where this step couples
        S.couple_body(method, float(dt))
        S.grid_settle(dvy)
        force = S.get_body()[1][0]
        if t > 0.5:  # measurement_start_time
            fx_sum += -float(np.float32(force[0]))
            fy_sum += float(np.float32(force[1]))
            samples += 1
        if i % 4 == 0:
            rows.append((t, x_i, float(force[0]), float(force[1]), float(force[2])))
    S.close()
    fx = float(f32(fx_sum / samples)) if samples else 0.0
    fy = float(f32(fy_sum / samples)) if samples else 0.0
    summary = {"avg_draft_fx_n": fx, "avg_vertical_fy_n": fy,
               "specific_draught_n_per_m": float(f32(fx) / (f32(2.0) * f32(he[2]))), "samples": samples}
    return np.array(rows, np.float64), summary


def write_metrics_csv(path: str, method_name: str, summary: dict) -> None:
    """The run summary line the harness writes beside the time series
    (bench_compare_all.cpp:160-176)."""
    s = summary
    with open(path, "w") as f:
        f.write(",".join(("method",) + METRIC_COLUMNS) + "\n")
        f.write("%s,%.6e,%.6e,%.6e,%.3f,%.6e,%.6e,%.4f,%.6f\n" % (
            method_name, s["max_abs_total_p_y"], s["avg_abs_total_p_y"], s["max_abs_energy_drift"],
            s["step_time_avg_us"], s["max_penetration_m"], s["avg_penetration_m"], s["dwell_fraction"],
            s["ratio_p_y_vs_M0"]))


def write_rebound_csv(path: str, method_name: str, summary: dict) -> None:
    """The foot step's rebound summary (bench_compare_all.cpp:481-491)."""
    s = summary
    with open(path, "w") as f:
        f.write(",".join(("method",) + ReboundTracker.COLUMNS) + "\n")
        f.write("%s,%.6e,%.6e,%.6f\n" % (method_name, s["v_down_peak"], s["v_up_peak"], s["rebound_coefficient_e"]))


def write_timeseries_csv(path: str, rows: dict) -> None:
    """FrameExporter's time-series format (frame_exporter.h:25-27, 44-57)."""
    with open(path, "w") as f:
        f.write(",".join(COLUMNS) + "\n")
        for i in range(len(rows["t"])):
            f.write("%.6f,%.6e,%.6e,%.6e,%.6e,%.8e,%.8e,%.8e,%.6e,%d\n" % (
                rows["t"][i], rows["box_x"][i], rows["box_y"][i], rows["box_z"][i], rows["box_vy"][i],
                rows["p_body_y"][i], rows["p_grid_y"][i], rows["p_sum_y"][i], rows["KE"][i],
                int(rows["nodes_inside"][i])))
