"""The 3601 x 3601 tile at 30 m, LoS 40 km (1333 steps per band), through
the C ABI.  The oracle checks band data of all 180 sectors, whole sectors
walked alone in three launch shapes, and the cumulative surfaces after
every 30 sectors up to the full 180-sector pass; the full run is also
checked as batched = plain; two sectors' ring events against the
oracle's records."""
import os
import sys

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def large(oracle):
    note("large-3601: DEM, engine bands")
    wl = demgen.large_3601()
    e = tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                       wl.observer_height, 0)
    e.upload(wl.elevations)
    e.compute_bands(0, 180)
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    yield wl, dem, e
    e.close()


@pytest.mark.parametrize("sector", [0, 45, 91, 179])
def test_large_band_data_bit_exact(large, oracle, sector):
    wl, dem, e = large
    note(f"large-3601 band data sector {sector}")
    want_d, want_x = oracle.band_data(dem, sector)
    got_d, got_x = e.get_band(sector)
    np.testing.assert_array_equal(got_d, want_d)
    np.testing.assert_array_equal(got_x.view(np.uint32), want_x.view(np.uint32))


SAMPLE = [7, 100, 179]   # 179 opens with a step at distance +0


@pytest.fixture(scope="module")
def sample_truth(large, oracle):
    wl, dem, _ = large
    deltas, dists = oracle.band_data_all(dem, SAMPLE)
    note(f"large-3601 oracle walk, sectors {SAMPLE}")
    want, _ = oracle.walk_sectors(dem, wl.elevations.astype(np.float32), wl.observer_height,
                                  deltas, dists, n_threads=THREADS)
    return want


@pytest.mark.parametrize("shape", [(64, 4, True, True), (128, 8, False, True),
                                   (256, 8, False, False)])
def test_large_sample_sectors_bit_exact(large, sample_truth, shape):
    """Sectors 7, 100 and 179 walked alone (one launch each) against the
    oracle."""
    wl, dem, e = large
    block, unroll, pair, two = shape
    sectors = SAMPLE
    want = sample_truth
    note(f"large-3601 engine, shape {shape}")
    e.set_walk_shape(block, unroll, True, both_directions=pair, two_points=two)
    e.reset_surfaces()
    for s in sectors:
        e.run(s, 1, "batched")
    np.testing.assert_array_equal(e.surfaces().view(np.uint32), want.view(np.uint32))


def test_large_full_run_batched_equals_plain(large):
    wl, dem, e = large
    note("large-3601 full plain")
    e.reset_surfaces()
    e.run(0, 180, "plain")
    plain = e.surfaces()
    note("large-3601 full batched, two points per lane")
    e.set_walk_shape(64, 4, True, both_directions=True, two_points=True)
    e.reset_surfaces()
    e.run(0, 180, "batched")
    np.testing.assert_array_equal(e.surfaces().view(np.uint32), plain.view(np.uint32))
    assert e.short_divide() == (True, 180)
    assert np.all(plain > 0)


# ---- all 180 sectors against the oracle ------------------------------------
GROUP = 30
_prefix = {}   # end sector -> oracle cumulative surfaces after sectors [0, end)


@pytest.fixture(scope="module")
def large_all_bands(large, oracle):
    """The oracle's band data for all 180 sectors (one sector per task)."""
    from concurrent.futures import ThreadPoolExecutor
    wl, dem, _ = large
    note("large-3601 oracle band data, 180 sectors")
    with ThreadPoolExecutor(THREADS) as ex:
        parts = list(ex.map(lambda s: oracle.band_data(dem, s), range(180)))
    return np.stack([p[0] for p in parts]), np.stack([p[1] for p in parts])


def test_large_band_data_all_sectors(large, large_all_bands):
    wl, dem, e = large
    deltas, dists = large_all_bands
    for s in range(180):
        got_d, got_x = e.get_band(s)
        np.testing.assert_array_equal(got_d, deltas[s], err_msg=f"deltas, sector {s}")
        np.testing.assert_array_equal(got_x.view(np.uint32), dists[s].view(np.uint32),
                                      err_msg=f"distances, sector {s}")


def _oracle_prefix(oracle, wl, dem, bands, end):
    """Oracle cumulative surfaces after sectors [0, end), extended from the
    longest prefix already walked (each prefix is a copy)."""
    if end in _prefix:
        return _prefix[end]
    start = max([k for k in _prefix if k < end], default=0)
    cum = _prefix[start].copy() if start else np.zeros(dem.computable_points_count, np.float32)
    deltas, dists = bands
    note(f"large-3601 oracle walk, sectors {start}..{end - 1}")
    oracle.walk_sectors(dem, wl.elevations.astype(np.float32), wl.observer_height,
                        deltas[start:end], dists[start:end], n_threads=THREADS, cumulative=cum)
    _prefix[end] = cum
    return cum


@pytest.mark.parametrize("end", range(GROUP, 181, GROUP))
def test_large_all_sectors_equal_oracle(large, large_all_bands, oracle, end):
    """The batched walk over sectors [0, end) in one launch, in the
    directional block order the tuner picks for this tile, equals the
    oracle's cumulative surfaces after the same sectors (end = 180: the
    whole pass, compute.rs:101-115)."""
    wl, dem, e = large
    e.set_walk_shape(64, 8, 2, both_directions=True, two_points=1)
    e.reset_surfaces()
    e.run(0, end, "batched")
    got = e.surfaces()
    want = _oracle_prefix(oracle, wl, dem, large_all_bands, end)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32),
                                  err_msg=f"sectors 0..{end - 1}")


@pytest.mark.parametrize("sector", [100, 179])
def test_large_ring_events_equal_oracle(large, large_all_bands, oracle, sector):
    """One sector's ring events at 3601 x 3601 (1 740 978 bands of 1333
    steps; sector 179 opens with a step at distance +0): the compact export
    equals the oracle's fixed-slot records (kernel.rs:141-243) band for band
    -- count slot, PoV slot, and the alternating close/open DEM ids.  No band
    fills its R slots here, so the oracle's records hold every event."""
    wl, dem, e = large
    R = oracle.ring_count_per_band(wl.rings_per_km, wl.max_line_of_sight)
    note(f"large-3601 ring events sector {sector}: engine export")
    counts, events = e.compact_rings(sector)
    assert counts.size == 2 * dem.computable_points_count
    assert int(counts.max()) < R, "a band overflows its R slots"
    deltas, dists = large_all_bands
    note(f"large-3601 ring events sector {sector}: oracle records")
    _, want = oracle.walk_sectors(dem, wl.elevations.astype(np.float32), wl.observer_height,
                                  deltas[sector:sector + 1], dists[sector:sector + 1],
                                  rings_per_band=R, rings=True, n_threads=THREADS)
    want = want[0].reshape(-1, R)
    np.testing.assert_array_equal(want[:, 0], counts)
    from tvs_amd import rings
    np.testing.assert_array_equal(want[:, 1], rings.band_povs(e.geometry))
    # Slots 2 .. count of every band, row after row, are the packed events.
    slots = np.arange(2, R)[None, :] <= (want[:, :1].astype(np.int64))
    np.testing.assert_array_equal(want[:, 2:][slots], events)
    assert not np.any(want[:, 2:][~slots])
