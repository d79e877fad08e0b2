"""Compact ring export (tvs_compact_rings) and viewshed reconstruction:
the reference's ASCII viewsheds (compute.rs:357-416), the fixed-slot
records, and the oracle's records, bit for bit."""
import os

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen, rings

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.mark.parametrize("case", [0, 1, 2])
def test_viewshed_golden_from_compact_rings(golden, case):
    v = golden["viewsheds"][case]
    t = golden["tvs"]
    print(f"compact rings golden {v['source']}", flush=True)
    with tvs_amd.Engine(9, 1.0, 3, 5000.0, 1.8) as e:
        e.upload(np.array(t[v["dem"]]["elevations"], np.int16))
        e.compute_bands(0, 180)
        sectors = [e.compact_rings(s) for s in range(180)]
        assert rings.ascii_viewshed(e.geometry, v["pov_id"], sectors) == v["rows"]


def test_compact_equals_records_and_oracle(oracle):
    wl = demgen.small_201(seed=4)
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    R = oracle.ring_count_per_band(wl.rings_per_km, wl.max_line_of_sight)
    picked = [0, 33, 90, 151]
    deltas, dists = oracle.band_data_all(dem, picked)
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height, 1) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        for i, s in enumerate(picked):
            print(f"compact rings sector {s}", flush=True)
            counts, events = e.compact_rings(s)
            assert np.all(counts % 2 == 0) and np.all(counts >= 2)
            e.run(s, 1, "batched")
            np.testing.assert_array_equal(rings.to_records(e.geometry, counts, events, R),
                                          e.rings(0))
            _, want = oracle.walk_sectors(dem, wl.elevations.astype(np.float32),
                                          wl.observer_height, deltas[i:i + 1], dists[i:i + 1],
                                          rings_per_band=R, rings=True)
            np.testing.assert_array_equal(rings.to_records(e.geometry, counts, events, R),
                                          want[0])


def test_compact_rings_have_no_slot_limit(oracle):
    """With R = 3 the fixed slots overflow; the compact form keeps every
    event and agrees with a roomy fixed-slot run."""
    wl = demgen.small_201(seed=5)
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, 0.5,
                        wl.observer_height, 1) as small_r, \
            tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, 500.0,
                           wl.observer_height, 1) as big_r:
        for e in (small_r, big_r):
            e.upload(wl.elevations)
            e.compute_bands(7, 1)
            e.run(7, 1, "batched")
        counts, events = small_r.compact_rings(7)
        assert counts.max() > small_r.rings_per_band     # would overflow R slots
        assert small_r.ring_overflow() > 0
        np.testing.assert_array_equal(
            rings.to_records(big_r.geometry, counts, events, big_r.rings_per_band),
            big_r.rings(0))
