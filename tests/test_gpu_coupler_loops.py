"""The reference's grid-only coupler loops on the device against the time
series the reference committed (research/mpm_rigid_coupling/outputs/csv/
{box_drop,rover_wheel,box_drop_dense,foot_step}_{M0,M1,M4.5a,M4.5b,M4.5}.csv, copied into
tests/golden/coupler_loops.npz by tests/golden/make_coupler_loops.py).

Each loop runs 720 steps (480 for the foot step) of: body gravity, the coupler on the seeded sand grid,
reaction into the body, TimeIntegrator::integrate, grid settling -- the
couplers, the integrator and the settling on the GPU (basements_amd.harness).
Bars: nodes_inside equal at every step; every logged float within the CSV's
own print precision (%.6e / %.8e: 7 significant digits), i.e. 1e-6 of the
column family's scale (positions: the largest |coordinate| of the run,
velocity: the largest |v_y|, momenta: the largest |p_y| logged, energy: the
largest KE).  The
reaction sums run in a different order than the reference's (a fixed-order
FP64 tree against hash-map order in FP32), which stays below that for the
body.  The logged grid momentum (p_grid_y, p_sum_y) is a sum over the ~360
persistent nodes that the reference takes in FP32 (hash-map order) and the
harness in FP64, and the nodes carry that reaction-order perturbation from
step to step: its bar is 5e-6 of scale plus the recursive-summation bound of
the FP32 sum, (n - 1) 2^-24 sum |v_y m| (Higham, Accuracy and Stability,
eq. 4.4).  Measured worst case: 2.0e-6 of scale (rover wheel M4.5b, step 475,
the wheel rebounding under the tangent blend), with the body's v_y there
within 6e-8 of scale and nodes_inside equal throughout."""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "coupler_loops.npz")
SCENES = ("box_drop", "rover_wheel", "box_drop_dense", "foot_step")
RUNS = tuple(f"{s}_{m}" for s in SCENES for m in ("M0", "M1", "M4.5a", "M4.5b", "M4.5"))


@pytest.mark.parametrize("run", RUNS)
def test_grid_only_loop_reproduces_the_committed_series(run):
    from basements_amd import harness

    ref = np.load(GOLDEN)
    cols = [str(c) for c in ref["columns"]]
    R = ref[run]
    scen, method = run.rsplit("_", 1)
    rows, _ = harness.run(harness.SCENARIOS[scen](), harness.method_id(method), frames=True)
    G = np.stack([rows[c] for c in cols], 1)
    assert G.shape == R.shape
    np.testing.assert_allclose(G[:, 0], R[:, 0], atol=1e-6)  # t, printed %.6f
    assert np.array_equal(G[:, 9], R[:, 9]), np.nonzero(G[:, 9] != R[:, 9])[0][:10]
    c = {k: cols.index(k) for k in cols}
    pos = np.abs(R[:, [c["box_x"], c["box_y"], c["box_z"]]]).max()
    vel = np.abs(R[:, c["box_vy"]]).max()
    mom = np.abs(R[:, [c["p_body_y"], c["p_grid_y"], c["p_sum_y"]]]).max()
    families = {"box_x": pos, "box_y": pos, "box_z": pos, "box_vy": vel, "p_body_y": mom, "p_grid_y": mom,
                "p_sum_y": mom, "KE": np.abs(R[:, c["KE"]]).max()}
    fp32_sum = np.maximum(rows["n_grid"] - 1, 0) * 2.0 ** -24 * rows["p_grid_abs"]
    for name, scale in families.items():
        d = np.abs(G[:, c[name]] - R[:, c[name]])
        grid = name in ("p_grid_y", "p_sum_y")
        bar = (5e-6 * scale + fp32_sum) if grid else 1e-6 * scale
        k = int(np.argmax(d - bar))
        assert d[k] <= bar[k] if np.ndim(bar) else d[k] <= bar, (run, name, k, G[k, c[name]], R[k, c[name]], scale)
    # the run summary (RunMetrics, metrics.h) against the committed <run>_metrics.csv: maxima are single
    # logged values (print precision of %.6e plus the series bars above), the means and the energy drift
    # FP32 accumulations over 720 steps (720 * 2^-24 ~ 4e-5 relative), the dwell fraction a count of
    # steps (one step = 1/720) printed with %.4f; the wall-clock step time is not compared
    got = rows["metrics"]
    want = dict(zip([str(x) for x in ref["metric_columns"]], ref[run + "_metrics"]))
    for name, rel in (("max_abs_total_p_y", 1e-5), ("max_penetration_m", 1e-5), ("avg_abs_total_p_y", 1e-4),
                      ("avg_penetration_m", 1e-4), ("max_abs_energy_drift", 1e-4)):
        assert abs(got[name] - want[name]) <= rel * abs(want[name]) + 1e-7, (run, name, got[name], want[name])
    assert abs(got["dwell_fraction"] - want["dwell_fraction"]) <= 1.5 / 720 + 5e-5, (run, got, want)
    # the foot step's rebound summary (ReboundTracker, <run>_rebound.csv): its peaks are logged v_y values
    # (the velocity bar, plus %.6e print precision), e printed with %.6f
    if run + "_rebound" in ref.files:
        rb, want_rb = rows["rebound"], ref[run + "_rebound"]
        for j, name in enumerate(("v_down_peak", "v_up_peak")):
            assert abs(rb[name] - want_rb[j]) <= 1e-6 * vel + 5e-7 * abs(want_rb[j]), (run, name, rb, want_rb)
        assert abs(rb["rebound_coefficient_e"] - want_rb[2]) <= 1e-6, (run, rb, want_rb)
    # grid snapshots (FrameExporter::log_frame; box drop only, every 8th committed frame): the same nodes
    # (mass > 0) at the same centres, velocities within the grid-momentum bar above (5e-6 of the run's
    # velocity scale: nodes at rest carry ~1e-11 m/s of FP32 noise there), masses to print precision
    if run + "_frames" not in ref.files:
        return
    F = ref[run + "_frames"].astype(np.float64)
    vmax = max(np.abs(F[:, 4:7]).max(), np.abs(R[:, c["box_vy"]]).max())
    for k in np.unique(F[:, 0]).astype(int):
        want_n = F[F[:, 0] == k][:, 1:]
        got_n = rows["frames"][k][3].astype(np.float64)
        assert got_n.shape == want_n.shape, (run, k, got_n.shape, want_n.shape)
        order = lambda a: np.lexsort((a[:, 2], a[:, 1], a[:, 0]))  # noqa: E731
        g, w = got_n[order(got_n)], want_n[order(want_n)]
        np.testing.assert_allclose(g[:, :3], w[:, :3], rtol=1e-6, atol=1e-7, err_msg=f"{run} frame {k} centres")
        np.testing.assert_allclose(g[:, 3:6], w[:, 3:6], rtol=1e-6, atol=5e-6 * vmax,
                                   err_msg=f"{run} frame {k} velocities")
        np.testing.assert_allclose(g[:, 6], w[:, 6], rtol=1e-6, atol=1e-9, err_msg=f"{run} frame {k} masses")



@pytest.mark.parametrize("method", ["M4.5a", "M4.5b", "M4.5"])
def test_m45_couplers_in_the_coupled_substep_batch_equals_alone(method):
    """The M4.5 methods as the coupler of the particle substep
    (bm_mpm_step_coupled) and of a batch scene: the box presses into the soil
    (nodes inside, finite reaction and state), and the scene in a batch
    evolves as it does alone (the batch tests' 1e-4 bar)."""
    import math

    from basements_amd import BODY_DYNAMIC, M45, M45A, M45B, SHAPE_BOX, Batch, Body, SceneDesc
    from _util import assert_close_field, gpu_solver, params_from, random_scene

    m = {"M4.5a": M45A, "M4.5b": M45B, "M4.5": M45}[method]
    P = random_scene(41, dx=1 / 32, cells_lo=(6, 0, 6), cells_hi=(16, 8, 16), vel=0.1, strain=0.01, cgrad=0.5)
    prm = params_from(E=3.5e5, alpha=0.386)
    q = (math.cos(0.1), 0.0, math.sin(0.1), 0.0)
    body = Body(position=(0.35, 0.27, 0.35), orientation=q, linear_velocity=(0.0, -0.3, 0.0),
                angular_velocity=(0.0, 0.2, 0.0), inv_mass=0.5, linear_damping=0.0, angular_damping=0.0,
                applied_force=(0.0, -9.81 * 2.0, 0.0), half_extents=(0.08, 0.05, 0.06), friction=0.5, mass=2.0,
                shape=SHAPE_BOX, integrate=BODY_DYNAMIC)
    sd = SceneDesc(dx=1 / 32, origin=(-4, -4, -4), dims=(32, 20, 32), E=prm.E, nu=prm.nu, rho=prm.rho,
                   alpha=prm.alpha, mu_floor=prm.mu_floor, coupler=m)
    k, dt = 12, 2e-4
    S = gpu_solver(P, sd.dx, sd.origin, sd.dims, prm, mode=1)
    S.set_body(body)
    S.step_coupled(m, dt, k)
    alone = S.download()
    b_alone, (force, torque, inside) = S.get_body()
    assert inside > 0 and np.isfinite(force).all() and np.isfinite(torque).all()
    for key in ("x", "v", "F", "C"):
        assert np.isfinite(alone[key]).all(), key
    B = Batch(max_particles=100_000, max_blocks=20_000, max_scenes=2)
    sid = B.add_scene(sd, P, body)
    B.step(dt, k)
    got = B.download_scene(sid)
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], alone[key], 1e-4, f"{method} batch {key}")
    assert_close_field(B.get_body(sid)[0].position, b_alone.position, 1e-5, f"{method} body position")


@pytest.mark.parametrize("method", ["M0", "M1", "M4.5a", "M4.5b"])
def test_tine_drag_reproduces_the_committed_series(method):
    """run_tine_drag (bench_compare_all.cpp:215-292): a kinematic tine dragged
    through the soil bed, the reaction logged every 4th step and averaged
    after 0.5 s, against tine_drag_<method>.csv and its summary.  The tine's
    path is prescribed, so its position matches to print precision; the
    reaction comes from the persistent grid, which carries the reaction-order
    perturbation from step to step: the same 5e-6-of-scale bar as the grid
    momentum of the other loops."""
    from basements_amd import harness

    ref = np.load(GOLDEN)
    R = ref[f"tine_drag_{method}"]
    rows, summary = harness.run_tine_drag(harness.method_id(method))
    assert rows.shape == R.shape
    np.testing.assert_allclose(rows[:, 0], R[:, 0], atol=5e-5)  # t, printed %.4f
    np.testing.assert_allclose(rows[:, 1], R[:, 1], rtol=1e-6, atol=1e-7)
    fscale = max(np.abs(R[:, 2:5]).max(), 1e-30)
    np.testing.assert_allclose(rows[:, 2:5], R[:, 2:5], rtol=1e-6, atol=5e-6 * fscale, err_msg=method)
    want = dict(zip(("avg_draft_fx_n", "avg_vertical_fy_n", "specific_draught_n_per_m", "samples"),
                    ref[f"tine_drag_{method}_summary"]))
    assert summary["samples"] == int(want["samples"])
    for k in ("avg_draft_fx_n", "avg_vertical_fy_n", "specific_draught_n_per_m"):
        assert abs(summary[k] - want[k]) <= 1e-5 * abs(want[k]) + 5e-6 * fscale, (method, k, summary[k], want[k])
