"""The configured workloads end to end against the oracle, bit for bit.

* main-1201 (1201 x 1201 at 90 m, LoS 36 km): band data of all 180 sectors
  and the 180-sector TVS of the batched walk in the default shape
  (compute.rs:101-115, kernel.rs:71-256).
* tiles-64 (64 seeded tiles, sides 257-1201, 30 or 90 m): one batch, tuned
  as the benchmark tunes it, replayed from its hipGraph; every tile's
  180-sector TVS.
* ring capture where R is small: a 257-point tile at 30 m keeps R = 12 ring
  slots per band (compute.rs:94-98), so bands overflow (kernel.rs:141).
"""
import functools
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen

pytestmark = pytest.mark.gpu
THREADS = min(16, len(os.sched_getaffinity(0)))
TILE_GROUPS = 8


def oracle_bands(oracle, dem):
    """The oracle's band data for all 180 sectors, one sector per task."""
    with ThreadPoolExecutor(THREADS) as ex:
        parts = list(ex.map(lambda s: oracle.band_data(dem, s), range(180)))
    return np.stack([p[0] for p in parts]), np.stack([p[1] for p in parts])


def test_main_1201_all_sectors_equal_oracle(oracle):
    wl = demgen.main_1201()
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    deltas, dists = oracle_bands(oracle, dem)
    print("main-1201: 180 sectors on the device", flush=True)
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        for s in range(180):
            d, x = e.get_band(s)
            np.testing.assert_array_equal(d, deltas[s], err_msg=f"deltas, sector {s}")
            np.testing.assert_array_equal(x.view(np.uint32), dists[s].view(np.uint32),
                                          err_msg=f"distances, sector {s}")
        e.run(0, 180, "batched")
        got = e.surfaces()
        got64 = e.surfaces_f64()
    print("main-1201: 180 sectors on the oracle", flush=True)
    want, _ = oracle.walk_sectors(dem, wl.elevations.astype(np.float32), wl.observer_height,
                                  deltas, dists, n_threads=THREADS)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # The f64 per-point sums of the same band surfaces stay close to the f32 ones.
    assert np.max(np.abs(got64 - want) / got64) < 1e-5


@pytest.fixture(scope="module")
def tiles64_device():
    """The 64 tiles in one batch, the shape picked by the batch tuner, two
    graph replays (an eager reset between them): every tile's surfaces."""
    tiles = demgen.tiles_64()
    print(f"tiles-64: {len(tiles)} tiles, batch tune and graph replays", flush=True)
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.compute_bands()
        shape, _ = b.autotune(reps=1)
        print(f"tiles-64 tuned shape {shape}", flush=True)
        b.run(True)
        first = [b.surfaces(i) for i in range(len(tiles))]
        b.reset_surfaces()
        b.run(True)
        b.sync()
        for i in range(len(tiles)):
            np.testing.assert_array_equal(b.surfaces(i), first[i], err_msg=f"replay, tile {i}")
    return tiles, first


@functools.lru_cache(maxsize=None)
def _bands_for(oracle, width, spacing, los):
    dem = oracle.dem_new(width, spacing, los)
    return oracle_bands(oracle, dem)


@pytest.mark.parametrize("group", range(TILE_GROUPS))
def test_tiles64_equal_oracle(oracle, tiles64_device, group):
    tiles, got = tiles64_device
    per = len(tiles) // TILE_GROUPS
    for i in range(group * per, (group + 1) * per):
        t = tiles[i]
        print(f"tiles-64 tile {i}: {t.width} x {t.width} at {t.spacing:g} m", flush=True)
        dem = oracle.dem_new(t.width, t.spacing, t.max_line_of_sight)
        # Band data depends on (width, band size) only; L = w div 3 here.
        deltas, dists = _bands_for(oracle, t.width, t.spacing, t.max_line_of_sight)
        want, _ = oracle.walk_sectors(dem, t.elevations.astype(np.float32), t.observer_height,
                                      deltas, dists, n_threads=THREADS)
        np.testing.assert_array_equal(got[i].view(np.uint32), want.view(np.uint32),
                                      err_msg=f"tile {i}")


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_ring_capture_where_r_is_small(oracle, method):
    """R = 12 slots per band on a 257-point tile at 30 m (LoS 2550 m, 5
    rings/km).  The reference writes a band's events past its R slots into
    the next band's (kernel.rs:141); the engine drops and counts them.  So:
    the dropped count equals the oracle's spill count, every band's count
    slot (kernel.rs:243) equals the oracle's, every band's slots up to
    min(count, R - 1) -- all of a band that did not overflow -- equal the
    oracle's, and the slots past them are zero."""
    w, spacing = 257, 30.0
    los = (w // 3) * int(spacing)
    elev = demgen.fractal_dem(w, spacing, 2.4, 4000.0, 1007).reshape(-1)
    dem = oracle.dem_new(w, spacing, los)
    R = oracle.ring_count_per_band(5.0, los)
    assert R == 12
    checked = overflowed = 0
    with tvs_amd.Engine(w, spacing, los, 5.0, 1.65, 1) as e:
        e.upload(elev)
        e.compute_bands(0, 180)
        for s in range(0, 180, 15):
            try:
                want_cum, ring, spill = oracle.compute_run(
                    dem, elev.astype(np.float32), 1.65, 5.0, first_sector=s, sector_count=1,
                    rings=True)
            except oracle.OracleError:
                continue    # the last band spills past the buffer: the reference panics
            print(f"small R: sector {s}, {method}, oracle spill {spill}", flush=True)
            e.reset_surfaces()
            e.run(s, 1, method)
            got = e.rings(0).reshape(-1, R)
            want = ring[0].reshape(-1, R)
            assert e.ring_overflow() == spill
            np.testing.assert_array_equal(got[:, 0], want[:, 0], err_msg=f"counts, sector {s}")
            counts = got[:, 0].astype(np.int64)
            upto = np.minimum(counts, R - 1)
            slot = np.arange(R)[None, :]
            own = (slot >= 1) & (slot <= upto[:, None])
            np.testing.assert_array_equal(got[own], want[own], err_msg=f"events, sector {s}")
            assert np.all(got[(slot > upto[:, None])] == 0)
            np.testing.assert_array_equal(e.surfaces(), want_cum, err_msg=f"surfaces, sector {s}")
            checked += 1
            overflowed += int(np.count_nonzero(counts > R - 1))
    assert checked >= 6 and overflowed > 0, (checked, overflowed)
