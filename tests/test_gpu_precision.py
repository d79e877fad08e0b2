"""Precision form on the 1201 x 1201 workload: int16 heights in HBM, f32
angles, f64 per-point sums over the 180 sectors, compared with a full-FP64
walk of the oracle (same band data; height difference, angle, max-scan and
surface sum in double).  Reports the visibility flips of the f32 walk and
the TVS relative-error distribution (max, p99), and writes them to
precision_main-1201.json in the folder named by TVS_REPORT_DIR when that
variable is set."""
import json
import os
import sys

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def test_precision_report_main_1201(oracle):
    wl = demgen.main_1201(seed=0)
    print("precision: engine int16 batched, 180 sectors", flush=True)
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        e.run(0, 180, "batched")
        tvs32 = e.surfaces().astype(np.float64)
        tvs_sum64 = e.surfaces_f64()
        deltas = np.stack([e.get_band(s)[0] for s in range(180)])
        dists = np.stack([e.get_band(s)[1] for s in range(180)])
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    print(f"precision: oracle f64 walk, {THREADS} threads", flush=True)
    tvs64, flips, steps = oracle.walk_sectors_f64(dem, wl.elevations.astype(np.float32),
                                                  wl.observer_height, deltas, dists,
                                                  n_threads=THREADS)
    rel32 = np.abs(tvs32 - tvs64) / tvs64
    rel64 = np.abs(tvs_sum64 - tvs64) / tvs64
    report = {
        "workload": wl.name,
        "point_steps": steps,
        "visibility_flips": flips,
        "flip_rate": flips / steps,
        "tvs_f32_vs_fp64_oracle": {"max": float(rel32.max()),
                                   "p99": float(np.percentile(rel32, 99)),
                                   "median": float(np.median(rel32))},
        "tvs_f64sum_vs_fp64_oracle": {"max": float(rel64.max()),
                                      "p99": float(np.percentile(rel64, 99)),
                                      "median": float(np.median(rel64))},
        "points_over_1e-4_f32": int((rel32 > 1e-4).sum()),
        "points": int(rel32.size),
    }
    print(json.dumps(report), flush=True)
    out = os.environ.get("TVS_REPORT_DIR", "")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "precision_main-1201.json"), "w") as f:
            json.dump(report, f, indent=1)
    sys.stdout.flush()
    assert steps == 180 * 2 * dem.computable_points_count * dem.n_deltas
    assert flips < steps * 1e-3
    assert report["tvs_f32_vs_fp64_oracle"]["p99"] < 1e-3
