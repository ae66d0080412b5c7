"""The reference's angle-of-repose scenario, for the tests: the particle column
of seed_repose_column and the pile metrics of measure_pile
(research/mpm_rigid_coupling/scenarios/scenario_repose_pile.h:58-172), in FP32
as the reference computes them."""
from __future__ import annotations

import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "repose_pile.json")
f32 = np.float32


def reference_run():
    with open(GOLDEN) as f:
        return json.load(f)


def column(sc):
    """Particle positions of the column, in add_particle order (i, j, k)."""
    nx, ny, nz = sc["column"]
    sp = f32(sc["spacing"])
    base = np.array(sc["base"], np.float32)
    pos = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                off = np.array([(f32(i) - f32(nx) * f32(0.5) + f32(0.5)) * sp, f32(j) * sp,
                                (f32(k) - f32(nz) * f32(0.5) + f32(0.5)) * sp], np.float32)
                pos.append(base + off)
    return np.asarray(pos, np.float32)


def measure(x, v, sc):
    """measure_pile: height, footprint radius, slope angle and geometric angle
    (degrees), particles at rest (|v| < 0.01 m/s)."""
    x = np.asarray(x, np.float32)
    v = np.asarray(v, np.float32)
    axis = np.array(sc["base"], np.float32)
    ground = f32(0.0)
    ymax = x[:, 1].max()
    speed = np.sqrt((v.astype(np.float64) ** 2).sum(1))
    at_rest = int((speed < 0.01).sum())
    height = float(ymax - ground)
    nbins = int(sc["bins"])
    r_max = f32(1.5) * (f32(sc["column"][0]) * f32(sc["spacing"]))
    bin_dr = r_max / f32(nbins)
    sy = np.full(nbins, -1.0, np.float32)
    sr = np.zeros(nbins, np.float32)
    r_outer = f32(0.0)
    for p in x:
        dxh, dzh = p[0] - axis[0], p[2] - axis[2]
        r = np.sqrt(f32(dxh * dxh + dzh * dzh))
        r_outer = max(r_outer, r)
        b = int(r / bin_dr)
        if 0 <= b < nbins and p[1] > sy[b]:
            sy[b] = p[1]
            sr[b] = r
    crest = int(np.argmax(sy))
    pts = [(sr[b], sy[b]) for b in range(crest, nbins) if sy[b] >= 0.0 and sy[b] >= ground + f32(0.005)]
    slope_deg = 0.0
    if len(pts) >= 2:
        r_ = np.array([p[0] for p in pts], np.float64)
        y_ = np.array([p[1] for p in pts], np.float64)
        n = len(pts)
        den = n * (r_ * r_).sum() - r_.sum() ** 2
        if abs(den) > 1e-9:
            slope = (n * (r_ * y_).sum() - r_.sum() * y_.sum()) / den
            slope_deg = math.degrees(math.atan(abs(slope)))
    geom_deg = math.degrees(math.atan(height / float(r_outer))) if r_outer > 1e-6 else 0.0
    return {"height": height, "radius": float(r_outer), "angle_slope": slope_deg, "angle_geom": geom_deg,
            "at_rest": at_rest}
