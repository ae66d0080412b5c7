"""Seeded synthetic soil scenes (the benchmark and test inputs).

Positions are drawn from a counter-based generator (Philox keyed by the seed)
so the same (seed, particle index) always yields the same particle on every
host.  Particles sit one per half-cell sub-cube, jittered uniformly inside it
(8 per cell), with v = 0, F = I, C = 0 and V0 = m / rho as
MPMSolver::add_particle sets them (mpm_solver.h:23-34 in the reference).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.Philox(key=int(seed)))


def soil_block(dx, cell_lo, cell_hi, seed=0, rho=1600.0, ppc_axis=2):
    """Particles filling cells [cell_lo, cell_hi) (integer cell coords of a dx grid).

    ppc_axis sub-cubes per axis per cell (2 -> 8 particles per cell).
    Returns the particle dict (float32 arrays) used by Solver.upload.
    """
    lo = np.asarray(cell_lo, np.int64)
    hi = np.asarray(cell_hi, np.int64)
    sub = (hi - lo) * ppc_axis
    n = int(np.prod(sub))
    idx = np.indices(tuple(int(s) for s in sub)).reshape(3, -1).T.astype(np.float64)
    h = dx / ppc_axis
    jit = _rng(seed).random((n, 3))
    x = (lo[None, :] * dx + (idx + jit) * h).astype(np.float32)
    # particle mass = rho * (sub-cube volume), rounded once to float32
    m = np.float32(rho * h ** 3)
    return make_particles(x, mass=m, rho=rho)


def make_particles(x, v=None, mass=1.0, rho=1500.0, F=None, Cm=None):
    x = np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1, 3))
    n = x.shape[0]
    m = np.broadcast_to(np.asarray(mass, np.float32), (n,)).astype(np.float32)
    rho32 = np.float32(rho)
    V0 = (m / rho32).astype(np.float32) if rho32 > np.float32(1e-12) else np.ones(n, np.float32)
    return {
        "x": x,
        "v": np.zeros((n, 3), np.float32) if v is None else np.asarray(v, np.float32).reshape(n, 3),
        "F": np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)) if F is None else np.asarray(F, np.float32).reshape(n, 3, 3),
        "C": np.zeros((n, 3, 3), np.float32) if Cm is None else np.asarray(Cm, np.float32).reshape(n, 3, 3),
        "m": m,
        "V0": V0,
    }


def dp_alpha(friction_deg):
    """alpha = sqrt(2/3) * 2 sin(phi) / (3 - sin(phi))  (Klar 2016; mpm_solver.h:103-104)."""
    s = math.sin(math.radians(friction_deg))
    return math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


@dataclass
class Scene:
    name: str
    dx: float
    particles: dict
    origin: tuple
    dims: tuple
    E: float = 3.5e5
    nu: float = 0.3
    rho: float = 1600.0
    alpha: float = dp_alpha(35.0)
    mu_floor: float = 0.4
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1e-4
    substeps: int = 1000
    tool: dict = field(default_factory=dict)
    walls: int = 0  # slip-wall thickness in nodes on the box faces (0: floor only)

    @property
    def n(self):
        return int(self.particles["x"].shape[0])


def box_drop(seed=0, scale=1):
    """Soil block 0.5 x 0.25 x 0.5 m on a 64^3 grid (dx = 1/64 m), 8 ppc
    (131,072 particles at scale 1), rho 1600, E 3.5e5, nu 0.3, friction 35 deg,
    with a 0.1 m rigid cube 0.1 m above the surface (M0 Dirichlet).
    ``scale`` > 1 shrinks the block per axis for quick tests."""
    dx = 1.0 / 64.0
    lo = np.array([16, 0, 16])
    hi = np.array([48, 16, 48])
    if scale > 1:
        ext = (hi - lo) // scale
        lo = np.array([32 - ext[0] // 2, 0, 32 - ext[2] // 2])
        hi = lo + np.maximum(ext, 1)
    P = soil_block(dx, lo, hi, seed=seed, rho=1600.0)
    top = hi[1] * dx
    tool = {"kind": "box", "half": (0.05, 0.05, 0.05), "center": (0.5, top + 0.1 + 0.05, 0.5), "friction": 0.5,
            "method": 0}
    return Scene("box_drop", dx, P, origin=(-4, -4, -4), dims=(72, 72, 72), tool=tool, substeps=1000)


def wheel_bed(seed=0, friction_deg=None, cells=(120, 36, 120), scale=1):
    """Soil bed 1.0 x 0.3 x 1.0 m at 8 ppc on a 256^3-node grid.

    The stated particle count (~4.2M) fixes the spacing: dx = 1/120 m gives
    120 x 36 x 120 cells = 4,147,200 particles; the 256^3-node domain then
    spans 2.13 m per axis.  A cylinder wheel (r 0.1 m, width 0.08 m) rolls
    across at 0.2 m/s under a 50 N load through the M1 coupler.  The friction
    angle is drawn from U(25, 40) deg per seed unless given."""
    dx = 1.0 / 120.0
    c = np.array(cells) // scale
    if friction_deg is None:
        friction_deg = float(_rng(seed + 7919).uniform(25.0, 40.0))
    lo = np.array([68, 0, 68]) if scale == 1 else np.array([4, 0, 4])
    P = soil_block(dx, lo, lo + c, seed=seed, rho=1600.0)
    dims = (256, 256, 256) if scale == 1 else tuple(int(v) for v in (c[0] + 16, c[1] + 24, c[2] + 16))
    origin = (0, -4, 0)
    top = float(c[1] * dx)
    R = 0.1
    tool = {"kind": "cylinder", "radius": R, "half_len": 0.04,
            "center": (float((lo[0] + 0.2 * c[0]) * dx), top + R - 0.01, float((lo[2] + 0.5 * c[2]) * dx)),
            "speed": 0.2, "load": 50.0, "friction": 0.5, "method": 1}
    return Scene("wheel_bed", dx, P, origin=origin, dims=dims, alpha=dp_alpha(friction_deg), tool=tool,
                 substeps=2000)


def tool_scene(seed, n_target, grid_edge=None, name="tool", depth=None):
    """One soil-tool scene of the batched workload: a soil bed of about
    n_target particles (8 per cell, width = 2 x height) in a cubic box of
    grid_edge nodes (32..128; the smallest that holds the bed when None), a box
    or cylinder tool sunk into the bed top and dragged along x at a speed
    from U(0.05, 0.5) m/s (kinematic), friction angle from U(25, 40) deg,
    E log-uniform in [1e5, 1e6] Pa.  dt = 1e-4 s.  `depth`: how far the
    tool's bottom sits below the bed top (default two cells)."""
    r = _rng(1_000_003 + seed)
    cells = max(8.0, n_target / 8.0)
    H = max(3, int(round((cells / 4.0) ** (1.0 / 3.0))))
    W = max(4, int(round(math.sqrt(cells / H))))
    edges = (32, 48, 64, 96, 128)
    G = grid_edge or next((e for e in edges if e >= W + 16), 128)
    dx = 1.0 / G
    lo = ((G - W) // 2, 4, (G - W) // 2)
    hi = (lo[0] + W, 4 + H, lo[2] + W)
    P = soil_block(dx, lo, hi, seed=seed, rho=1600.0)
    E = float(10.0 ** r.uniform(5.0, 6.0))
    phi = float(r.uniform(25.0, 40.0))
    speed = float(r.uniform(0.05, 0.5))
    kind = "box" if r.random() < 0.5 else "cylinder"
    top = hi[1] * dx
    size = max(3, W // 6) * dx  # the tool spans a sixth of the bed
    sink = 2 * dx if depth is None else float(depth)
    # a small off-grid shift keeps every tool face off the node planes at
    # (i + 1/2) dx: a face on a node plane leaves that plane's nodes inside or
    # outside by rounding alone
    cx = (lo[0] + 0.25 * W + 0.0123) * dx
    cz = (lo[2] + 0.5 * W + 0.0137) * dx
    if kind == "box":
        he = (0.5 * size, 0.5 * size, 0.5 * size)
        tool = {"kind": "box", "half": he, "center": (cx, top + 0.5 * size - sink, cz), "speed": speed,
                "friction": 0.5, "mass": 1.0, "method": 0}
    else:
        R = 0.5 * size
        tool = {"kind": "cylinder", "radius": R, "half_len": 0.5 * size,
                "center": (cx, top + R - sink, cz), "speed": speed, "friction": 0.5, "load": 10.0, "method": 1}
    return Scene(name, dx, P, origin=(0, 0, 0), dims=(G, G, G), E=E, rho=1600.0, alpha=dp_alpha(phi), tool=tool,
                 substeps=500)


def batch_scenes(seed=0, count=64, n_lo=2e4, n_hi=5e5):
    """The batched workload: `count` tool scenes with particle counts
    log-uniform in [n_lo, n_hi]."""
    r = _rng(seed)
    ns = np.exp(r.uniform(math.log(n_lo), math.log(n_hi), size=count))
    return [tool_scene(seed * 1000 + i, float(n), name=f"tool{i}") for i, n in enumerate(ns)]


def serve_requests(seed=0, count=48, n_lo=5e4, n_hi=2e6, sub_lo=200, sub_hi=2000):
    """The served workload's requests: soil-tool rollouts with particle
    counts log-uniform in [n_lo, n_hi], tool depth U(0.02, 0.1) m, speed
    U(0.05, 0.5) m/s (drawn by tool_scene) and substeps uniform in
    {sub_lo..sub_hi}.  The bed sits in a soil bin: slip walls three nodes
    deep on the box faces keep a slumping bed inside its box over the long
    rollouts.  Returns [(Scene, substeps)]."""
    r = _rng(seed + 424_242)
    ns = np.exp(r.uniform(math.log(n_lo), math.log(n_hi), size=count))
    depths = r.uniform(0.02, 0.1, size=count)
    subs = r.integers(sub_lo, sub_hi + 1, size=count)
    out = []
    for i, (n, dp, k) in enumerate(zip(ns, depths, subs)):
        s = tool_scene(seed * 1000 + 500 + i, float(n), name=f"req{i}", depth=float(dp))
        s.walls = 3
        out.append((s, int(k)))
    return out


def poisson_arrivals(rate, count, seed=0):
    """Arrival times (s) of a Poisson stream of `rate` requests per second."""
    r = _rng(seed + 77_777)
    return np.cumsum(r.exponential(1.0 / rate, size=count))


def tool_body(tool):
    """The kinematic body of a tool scene (dragged along x at its speed)."""
    from basements_amd import BODY_KINEMATIC, SHAPE_BOX, SHAPE_CYLINDER, Body

    if tool["kind"] == "cylinder":
        R = tool["radius"]
        return Body(position=tool["center"], orientation=(math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0),
                    linear_velocity=(tool["speed"], 0.0, 0.0), angular_velocity=(0.0, 0.0, -tool["speed"] / R),
                    inv_mass=1.0 / tool["load"], linear_damping=0.0, angular_damping=0.0,
                    half_extents=(R, tool["half_len"], R), friction=tool["friction"], mass=tool["load"],
                    shape=SHAPE_CYLINDER, integrate=BODY_KINEMATIC)
    return Body(position=tool["center"], linear_velocity=(tool["speed"], 0.0, 0.0), inv_mass=1.0 / tool["mass"],
                linear_damping=0.0, angular_damping=0.0, half_extents=tool["half"], friction=tool["friction"],
                mass=tool["mass"], shape=SHAPE_BOX, integrate=BODY_KINEMATIC)


def dynamic_body(tool, g=9.81):
    """The scene's rigid body as the reference's scenarios build it: dynamic,
    undamped, gravity as an applied force, inertia assigned directly so the
    inverse inertia stays the identity (rigid_body.h:122,
    scenario_rover_wheel.h:95-133, scenario_box_drop.h:60-67).  A wheel rolls:
    local Y (axle) along world Z, v = the tool speed forward, w = -v/R."""
    from basements_amd import SHAPE_BOX, SHAPE_CYLINDER, Body

    if tool["kind"] == "cylinder":
        m = tool["load"] / g
        R = tool["radius"]
        return Body(position=tool["center"], orientation=(math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0),
                    linear_velocity=(tool["speed"], 0.0, 0.0), angular_velocity=(0.0, 0.0, -tool["speed"] / R),
                    inv_mass=1.0 / m, linear_damping=0.0, angular_damping=0.0, applied_force=(0.0, -g * m, 0.0),
                    half_extents=(R, tool["half_len"], R), friction=tool["friction"], mass=m, shape=SHAPE_CYLINDER)
    m = tool.get("mass", 1.0)
    return Body(position=tool["center"], inv_mass=1.0 / m, linear_damping=0.0, angular_damping=0.0,
                applied_force=(0.0, -g * m, 0.0), half_extents=tool["half"], friction=tool["friction"], mass=m,
                shape=SHAPE_BOX)


def tine_scene(seed=0, cells=(32, 4, 32), velocity=None):
    """The differentiable rollout: a soil bed of cells[0] x cells[1] x
    cells[2] cells (x, height, z) at dx = 1/32 m, 8 particles per cell
    (32 x 4 x 32 -> 32,768 particles), in a node box with 4 nodes of margin
    below and on the sides and room above; a 0.02 x 0.15 x 0.1 m tine
    (thickness along x, height, width) whose bottom sits 0.05 m below the
    bed top, dragged along x (velocity from U(0.05, 0.5) m/s by seed unless
    given) through M0; 400 substeps of 1e-4 s.  The tine is thinner than a
    cell, so it starts with the node plane x = (i + 1/2) dx just behind its
    front face; the plane stays inside while the tine moves less than
    1.95 cm (0.49 m/s over the 400 substeps).  The tine sits 1.1 mm off the
    bed's mid-plane in z and 0.7 mm deeper than 5 cm, so that no node inside
    it is equally far from two faces: at such a tie the M0 normal, and with
    it the loss, jumps under a small change of the tool's path, and central
    differences there measure the jump.  Material E = 3.5e5 Pa, nu 0.3,
    friction 35 deg.  The loss target is a seeded affine displacement field
    of about a millimetre: d(x0) = a + B (x0 - c).
    Returns a dict with particles, params (DiffParams), target, dx, origin,
    dims, dt, substeps."""
    from basements_amd import DiffParams

    dx = 1.0 / 32.0
    cx, cy, cz = (int(c) for c in cells)
    P = soil_block(dx, (0, 0, 0), (cx, cy, cz), seed=seed, rho=1600.0)
    r = _rng(seed + 31_337)
    v = float(r.uniform(0.05, 0.5)) if velocity is None else float(velocity)
    top = cy * dx
    i_plane = max(1, cx // 4)
    xc = (i_plane + 0.5) * dx - 0.0095
    prm = DiffParams(young_modulus=3.5e5, poisson_ratio=0.3, friction_deg=35.0, floor_friction=0.4,
                     tool=True, tool_center=(xc, top - 0.05 + 0.075 - 0.0007, 0.5 * cz * dx + 0.0011),
                     tool_half_extents=(0.01, 0.075, 0.05), tool_velocity=(v, 0.0, 0.0), tool_friction=0.5)
    x0 = P["x"].astype(np.float64)
    c = x0.mean(axis=0)
    a = r.normal(0.0, 1e-3, size=3)
    B = r.normal(0.0, 1e-2, size=(3, 3))
    target = (a[None, :] + (x0 - c) @ B.T).astype(np.float32)
    dims = (cx + 8, cy + 16, cz + 8)
    return {"particles": P, "params": prm, "target": target, "dx": dx, "origin": (-4, -4, -4), "dims": dims,
            "dt": 1e-4, "substeps": 400}
