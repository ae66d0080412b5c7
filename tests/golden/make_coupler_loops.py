"""Write tests/golden/coupler_loops.npz from the reference's committed outputs.

The reference's research harness committed the time series of its grid-only
coupler loops (research/mpm_rigid_coupling/outputs/csv/{box_drop,rover_wheel,
box_drop_dense,foot_step}_{M0,M1,M4.5a,M4.5b,M4.5}.csv, written by
FrameExporter::log_step, frame_exporter.h:44-57, from bench_compare_all.cpp:72-134,
:374-427, :505-560 and :614-689).  This copies their numbers into
one .npz (no pickles): key "<scenario>_<method>" -> (steps, 10) float64 in
the CSV's column order, and "<scenario>_<method>_metrics" -> the 8 numbers of
the run summary the harness wrote beside it (<run>_metrics.csv,
bench_compare_all.cpp:160-176; its method name dropped), and for the box
drop "<run>_frames" -> float32 rows (frame, x, y, z, vx, vy, vz, mass) of the
node rows of every FRAME_STRIDE-th grid snapshot of <run>_frames.csv
(FrameExporter::log_frame, frame_exporter.h:59-94), and for the foot step
"<run>_rebound" -> v_down_peak, v_up_peak, e of <run>_rebound.csv
(bench_compare_all.cpp:481-491).  tests/test_gpu_coupler_loops.py compares the device
loops with them.  Runs only where the reference tree is present.

    python tests/golden/make_coupler_loops.py
"""
from __future__ import annotations

import csv
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_CSV = "/root/reference/research/mpm_rigid_coupling/outputs/csv"
RUNS = tuple(f"{s}_{m}" for s in ("box_drop", "rover_wheel", "box_drop_dense", "foot_step")
             for m in ("M0", "M1", "M4.5a", "M4.5b", "M4.5"))
COLUMNS = ("t", "box_x", "box_y", "box_z", "box_vy", "p_body_y", "p_grid_y", "p_sum_y", "KE", "nodes_inside")
FRAME_STRIDE = 8
REBOUND = ("v_down_peak", "v_up_peak", "rebound_coefficient_e")
TINE_SUMMARY = ("avg_draft_fx_n", "avg_vertical_fy_n", "specific_draught_n_per_m", "samples")
METRICS = ("max_abs_total_p_y", "avg_abs_total_p_y", "max_abs_energy_drift", "step_time_avg_us",
           "max_penetration_m", "avg_penetration_m", "dwell_fraction", "ratio_p_y_vs_M0")


def main():
    out = {}
    for run in RUNS:
        with open(os.path.join(REF_CSV, run + ".csv")) as f:
            rows = list(csv.DictReader(f))
        out[run] = np.array([[float(r[c]) for c in COLUMNS] for r in rows], np.float64)
        with open(os.path.join(REF_CSV, run + "_metrics.csv")) as f:
            m = list(csv.DictReader(f))[0]
        out[run + "_metrics"] = np.array([float(m[c]) for c in METRICS], np.float64)
        if run.startswith("foot_step_"):
            with open(os.path.join(REF_CSV, run + "_rebound.csv")) as f:
                r = list(csv.DictReader(f))[0]
            out[run + "_rebound"] = np.array([float(r[c]) for c in REBOUND], np.float64)
        fpath = os.path.join(REF_CSV, run + "_frames.csv")
        if run.startswith("box_drop_M") and os.path.exists(fpath):
            with open(fpath) as f:
                fr = list(csv.reader(f))[1:]
            times = sorted({float(r[0]) for r in fr})
            keep = {t: k for k, t in enumerate(times) if k % FRAME_STRIDE == 0}
            out[run + "_frames"] = np.array([[keep[float(r[0])]] + [float(v) for v in r[3:10]]
                                             for r in fr if r[1] == "N" and float(r[0]) in keep], np.float32)
    for m in ("M0", "M1", "M4.5a", "M4.5b"):  # run_tine_drag, bench_compare_all.cpp:215-292
        with open(os.path.join(REF_CSV, f"tine_drag_{m}.csv")) as f:
            rows = list(csv.reader(f))[1:]
        out[f"tine_drag_{m}"] = np.array([[float(v) for v in r] for r in rows], np.float64)
        with open(os.path.join(REF_CSV, f"tine_drag_{m}_summary.csv")) as f:
            s = list(csv.DictReader(f))[0]
        out[f"tine_drag_{m}_summary"] = np.array([float(s[c]) for c in TINE_SUMMARY], np.float64)
    # the harness's M2 / M3 / M4 are skeletons delegating to M1 (couplers/m2_asflip.h, m3_mls_mpm.h,
    # m4_lagrange.h): their committed series equal M1's, which is what basements_amd.harness.method_id maps
    for s in ("box_drop", "rover_wheel", "box_drop_dense", "foot_step"):
        for m in ("M2", "M3", "M4"):
            with open(os.path.join(REF_CSV, f"{s}_{m}.csv")) as f:
                rows = list(csv.DictReader(f))
            alias = np.array([[float(r[c]) for c in COLUMNS] for r in rows], np.float64)
            assert np.array_equal(alias, out[f"{s}_M1"]), (s, m)
    out["columns"] = np.array(COLUMNS)
    out["metric_columns"] = np.array(METRICS)
    np.savez_compressed(os.path.join(HERE, "coupler_loops.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
