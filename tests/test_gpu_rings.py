"""Compact ring export (tvs_compact_rings) and viewshed reconstruction:
the reference's ASCII viewsheds (compute.rs:357-416), the fixed-slot
records, and the oracle's records, bit for bit."""
import os

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen, rings, _lib

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


def test_compact_rings_counts_kept_until_the_state_changes(oracle):
    """The counts of the last sector counted are kept for the next call on
    that sector (the size query and the events call count once) and
    dropped by an upload: after it, the export equals that of a fresh
    engine on the same state."""
    wl = demgen.small_201(seed=6)
    other = (wl.elevations[::-1].copy() + 7).astype(np.int16)

    def fresh(elev, sector):
        with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                            wl.observer_height) as f:
            f.upload(elev)
            f.compute_bands(sector, 1)
            return f.compact_rings(sector)

    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        first = e.compact_rings(12)
        again = e.compact_rings(12)
        for a, b in zip(first, again):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(first, fresh(wl.elevations, 12)):
            np.testing.assert_array_equal(a, b)
        e.compact_rings(13)                              # another sector in between
        for a, b in zip(e.compact_rings(12), first):
            np.testing.assert_array_equal(a, b)
        e.upload(other)                                  # upload: counted again
        got = e.compact_rings(12)
        want = fresh(other, 12)
        assert got[1].size != first[1].size or not np.array_equal(got[1], first[1])
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
        counts, events = got
        offsets = np.concatenate([[0], np.cumsum(counts.astype(np.int64) - 1)])
        assert offsets[-1] == events.size


def test_compact_rings_at_1201_equal_fixed_slots():
    """1201 x 1201: 321 602 bands, 315 tiles of the device scan of the
    offsets (two passes of its one-block tile scan); the compact export
    rebuilds the fixed-slot records of a capture run of the same sector."""
    wl = demgen.main_1201()
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height, 1) as e:
        e.upload(wl.elevations)
        e.compute_bands(46, 1)
        print("compact rings 1201 sector 46", flush=True)
        counts, events = e.compact_rings(46)
        assert counts.size == 321602 and np.all(counts % 2 == 0)
        e.run(46, 1, "batched")
        R = e.rings_per_band
        np.testing.assert_array_equal(rings.to_records(e.geometry, counts, events, R),
                                      e.rings(0))


def test_compact_rings_into_page_locked_memory():
    """tvs_host_alloc memory as the export's destination: the counts and
    events equal those copied into ordinary (pageable) arrays, and the
    allocation outlives the slices taken of it.  tvs_host_alloc(0) is NULL
    and tvs_host_free(NULL) succeeds."""
    L = _lib.load()
    assert not L.tvs_host_alloc(0)
    assert L.tvs_host_free(None) == 0
    wl = demgen.main_1201()
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height, 0) as e:
        e.upload(wl.elevations)
        e.compute_bands(46, 2)
        for s in (46, 47):
            print(f"compact rings 1201 sector {s}: page-locked and pageable", flush=True)
            pc, pe = e.compact_rings(s, pinned=True)
            qc, qe = e.compact_rings(s, pinned=False)
            assert pc.flags.writeable and pe.size == qe.size
            np.testing.assert_array_equal(pc, qc)
            np.testing.assert_array_equal(pe, qe)
    # The arrays are views of page-locked memory that they keep alive.
    import gc
    gc.collect()
    assert int(pc.sum()) == int(qc.sum()) and int(pe[-1]) == int(qe[-1])
