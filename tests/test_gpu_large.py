"""The 3601 x 3601 tile at 30 m, LoS 40 km (1333 steps per band), through
the C ABI.  The oracle checks band data and whole sectors on a sample; the
full 180-sector run is checked by the batched walk (tuned two-point shapes,
short divide) equalling the plain walk bit for bit."""
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
