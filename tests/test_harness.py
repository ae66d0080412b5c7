"""Host side of the grid-only coupler loops (basements_amd.harness): the
scenario constants against the reference's scenario headers, the committed
fixture's shape, and the time-series CSV format of FrameExporter::log_step
(frame_exporter.h:25-27, 44-57).  The device loops themselves are checked in
tests/test_gpu_coupler_loops.py."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "coupler_loops.npz")


def test_scenarios_follow_the_reference_headers():
    from basements_amd import harness

    b = harness.box_drop()
    # scenario_box_drop.h:14-40: 3 s at 1/240 s, 7 x 3 x 7 nodes of 0.01 kg at dx 0.1
    assert b.num_steps() == 720
    assert (b.extent_i, b.extent_k, b.layers_y, b.node_mass, b.dx) == (3, 3, 3, 0.01, 0.1)
    assert b.half_extents == (0.2, 0.2, 0.2) and b.position == (0.0, 0.55, 0.0)
    w = harness.rover_wheel()
    # scenario_rover_wheel.h:19-40: R 0.15, half length 0.05, 13 x 3 x 9 nodes, spin -10 rad/s about Z
    assert w.num_steps() == 720
    assert (w.extent_i, w.extent_k, w.layers_y) == (6, 4, 3)
    assert w.half_extents == (0.15, 0.05, 0.15) and w.angular_velocity == (0.0, 0.0, -10.0)
    # Quaternion::from_axis_angle((1,0,0), 1.5707963f): a quarter turn about X
    q = np.asarray(w.orientation)
    assert abs(np.dot(q, q) - 1.0) < 1e-6 and abs(q[0] - q[1]) < 1e-7 and q[2] == q[3] == 0.0
    # both patches and every pose they reach stay block-aligned with the reference's 4^3 blocks
    for sc in (b, w):
        assert all(o % 4 == 0 for o in sc.origin) and all(d % 4 == 0 for d in sc.dims)


def test_method_names_map_to_couplers():
    from basements_amd import M0, M1, M45, M45A, M45B, harness

    assert [harness.method_id(m) for m in ("M0", "M1", "M4.5a", "M4.5b", "M4.5")] == [M0, M1, M45A, M45B, M45]
    # the harness's M2 / M3 / M4 delegate to M1 (their committed series equal M1's,
    # checked by tests/golden/make_coupler_loops.py where the reference is present)
    assert {harness.method_id(m) for m in ("M2", "M3", "M4")} == {M1}


def test_seeded_patch_matches_seed_sand():
    from basements_amd import harness

    sc = harness.box_drop()
    mass, vel, active = harness._seed(sc)
    assert int(active.sum()) == 7 * 3 * 7
    assert np.all(mass[active] == np.float32(0.01)) and not vel.any()
    ii, jj, kk = np.nonzero(active)
    o = sc.origin
    assert (ii + o[0]).min() == -3 and (ii + o[0]).max() == 3
    assert (jj + o[1]).min() == 0 and (jj + o[1]).max() == 2


def test_fixture_holds_the_four_committed_series():
    ref = np.load(GOLDEN)
    assert [str(c) for c in ref["columns"]] == ["t", "box_x", "box_y", "box_z", "box_vy", "p_body_y", "p_grid_y",
                                                "p_sum_y", "KE", "nodes_inside"]
    runs = [(s, f"{s}_{m}") for s in ("box_drop", "rover_wheel", "box_drop_dense", "foot_step")
            for m in ("M0", "M1", "M4.5a", "M4.5b", "M4.5")]
    for s, run in runs:
        R = ref[run]
        steps = 480 if s == "foot_step" else 720  # 2 s / 3 s at 1/240 s
        assert R.shape == (steps, 10) and ref[run + "_metrics"].shape == (8,)
        np.testing.assert_allclose(R[:, 0], (np.arange(steps) + 1) / 240.0, atol=1e-6)
        # free fall before the first contact: v_y = -g t, as the integrator gives
        k = int(np.argmax(R[:, 9] > 0))
        n = min(k, 5)  # the printed t is rounded to 1e-6: use (i + 1) dt itself
        np.testing.assert_allclose(R[:n, 4], -9.81 * (np.arange(n) + 1) / 240.0, rtol=1e-5)


def test_fixture_holds_the_tine_drag_series():
    ref = np.load(GOLDEN)
    for m in ("M0", "M1", "M4.5a", "M4.5b"):
        R = ref[f"tine_drag_{m}"]
        # run_tine_drag logs every 4th of 480 steps: t = (4k + 1) / 240 s, the tine at -0.05 + 0.2 t
        assert R.shape == (120, 5)
        t = (4 * np.arange(120) + 1) / 240.0
        np.testing.assert_allclose(R[:, 0], t, atol=5e-5)
        np.testing.assert_allclose(R[:, 1], -0.05 + 0.2 * t, atol=1e-6)
        assert ref[f"tine_drag_{m}_summary"][3] == 360  # samples after 0.5 s


def test_metrics_writer_matches_the_reference_format(tmp_path):
    from basements_amd import harness

    m = harness.RunMetrics()
    m.observe(1.0, 2.0, 0.1, 0.2, 5.0)   # first step: energy reference
    m.observe(-3.0, 2.2, 0.05, 0.2, 7.0)  # drift 0.1; 0.05 < 0.3 * 0.2: no dwell
    s = m.summary()
    assert s["max_abs_total_p_y"] == 3.0 and s["avg_abs_total_p_y"] == 2.0
    assert abs(s["max_abs_energy_drift"] - 0.1) < 1e-6 and s["dwell_fraction"] == 0.5
    p = tmp_path / "m.csv"
    harness.write_metrics_csv(str(p), "M0", s)
    lines = p.read_text().splitlines()
    assert lines[0] == ("method,max_abs_total_p_y,avg_abs_total_p_y,max_abs_energy_drift,step_time_avg_us,"
                        "max_penetration_m,avg_penetration_m,dwell_fraction,ratio_p_y_vs_M0")
    assert lines[1].startswith("M0,3.000000e+00,2.000000e+00,1.000000e-01,6.000,1.000000e-01,") and \
        lines[1].endswith(",0.5000,1.000000")


def test_frames_writer_matches_the_reference_format(tmp_path):
    from basements_amd import harness

    nodes = np.array([[-0.25, 0.05, -0.25, 0.0, 0.0, 0.0, 0.01]], np.float32)
    p = tmp_path / "fr.csv"
    harness.write_frames_csv(str(p), [(1 / 240, (0.0, 0.5498297, 0.0), (0.0, -0.040875, 0.0), nodes)])
    lines = p.read_text().splitlines()
    assert lines[0] == "t,kind,id,x,y,z,vx,vy,vz,mass"
    # the first rows of research/mpm_rigid_coupling/outputs/csv/box_drop_M1_frames.csv
    assert lines[1] == "0.004167,B,0,0.000000e+00,5.498297e-01,0.000000e+00,0.000000e+00,-4.087500e-02," \
                       "0.000000e+00,0.000000e+00"
    assert lines[2] == "0.004167,N,0,-2.500000e-01,5.000000e-02,-2.500000e-01,0.000000e+00,0.000000e+00," \
                       "0.000000e+00,1.000000e-02"


def test_csv_writer_matches_the_reference_format(tmp_path):
    from basements_amd import harness

    rows = {k: np.zeros(2) for k in harness.COLUMNS}
    rows["t"][:] = (1 / 240, 2 / 240)
    rows["nodes_inside"][:] = (0, 3)
    p = tmp_path / "ts.csv"
    harness.write_timeseries_csv(str(p), rows)
    lines = p.read_text().splitlines()
    assert lines[0] == "t,box_x,box_y,box_z,box_vy,p_body_y,p_grid_y,p_sum_y,KE,nodes_inside"
    assert lines[1] == ("0.004167,0.000000e+00,0.000000e+00,0.000000e+00,0.000000e+00,0.00000000e+00,"
                        "0.00000000e+00,0.00000000e+00,0.000000e+00,0")
    assert lines[2].endswith(",3")


def test_rebound_tracker_replays_the_committed_foot_step_summaries():
    """ReboundTracker fed the committed foot-step series' body v_y (%.6e, the
    same FP32 values the reference observed, printed) gives the committed
    <run>_rebound.csv: v_down_peak the most negative v_y from 0.05 s on."""
    from basements_amd import harness

    ref = np.load(GOLDEN)
    dt = np.float32(1.0 / 240.0)
    for m in ("M0", "M1", "M4.5a", "M4.5b", "M4.5"):
        R = ref[f"foot_step_{m}"]
        tr = harness.ReboundTracker()
        for i in range(R.shape[0]):
            tr.observe(float(np.float32(i + 1) * dt), R[i, 4], 0.05)
        want = ref[f"foot_step_{m}_rebound"]
        got = tr.summary()
        assert np.float32(got["v_down_peak"]) == np.float32(want[0]), (m, got, want)
        assert got["v_up_peak"] == want[1] and abs(got["rebound_coefficient_e"] - want[2]) <= 5e-7, (m, got, want)


def test_rebound_tracker_phases():
    """The tracker's three phases on hand-made series: samples before the
    settle time are ignored, the up peak ends at the first non-positive v_y
    after it, e is v_up / |v_down| capped at 2, and 0 without a descent."""
    from basements_amd import harness

    def track(vs, settle=0.0):
        tr = harness.ReboundTracker()
        for k, v in enumerate(vs):
            tr.observe(0.01 * (k + 1), v, settle)
        return tr.summary()

    s = track([-1.0, -4.0, -2.0, 0.5, 1.0, 0.2, -0.1, 3.0])  # the late 3.0 is after the rebound ended
    assert (s["v_down_peak"], s["v_up_peak"], s["rebound_coefficient_e"]) == (-4.0, 1.0, 0.25)
    assert track([-9.0, -1.0, -2.0], settle=0.015)["v_down_peak"] == -2.0  # -9 falls before the settle time
    assert track([-1.0, 0.0, 5.0, 1.0])["rebound_coefficient_e"] == 2.0
    assert track([0.5, 1.0])["rebound_coefficient_e"] == 0.0
    assert track([-3.0, -1.0])["v_up_peak"] == 0.0


def test_rebound_writer_matches_the_reference_format(tmp_path):
    from basements_amd import harness

    p = tmp_path / "reb.csv"
    harness.write_rebound_csv(str(p), "M1", {"v_down_peak": -1.921124, "v_up_peak": 0.0, "rebound_coefficient_e": 0.0})
    # research/mpm_rigid_coupling/outputs/csv/foot_step_M1_rebound.csv
    assert p.read_text().splitlines() == ["method,v_down_peak,v_up_peak,rebound_coefficient_e",
                                          "M1,-1.921124e+00,0.000000e+00,0.000000"]
