"""HIP engine vs the oracle on seeded synthetic DEMs, through the C ABI.

Bit-exact: band data (sort order, deltas, distances), full Axes orders,
cumulative f32 surfaces of both walk methods, ring records.  At the
1201 x 1201 size the oracle checks band data and the ring events of all 180
sectors and the surfaces of a sample of sectors (test_gpu_configs.py holds
the full 180-sector TVS); the full run is also checked by size-independent
properties (the two methods agree bit for bit; the f64 sums agree with the
f32 ones within 1e-5 relative)."""
import os
import sys

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen, _lib
from tvs_amd import rings as rings_mod

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# --- 201 x 201, all 180 sectors -------------------------------------------

@pytest.fixture(scope="module", params=[0, 1])
def small(request, oracle):
    wl = demgen.small_201(seed=request.param)
    note(f"small-201 seed {request.param}: oracle band data + walk")
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    deltas, dists = oracle.band_data_all(dem, range(180))
    elev = wl.elevations.astype(np.float32)
    cum, _ = oracle.walk_sectors(dem, elev, wl.observer_height, deltas, dists,
                                 n_threads=THREADS)
    return wl, dem, deltas, dists, cum


def make_engine(wl, capture=0, rings_per_km=None):
    return tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight,
                          wl.rings_per_km if rings_per_km is None else rings_per_km,
                          wl.observer_height, capture)


@pytest.mark.parametrize("band_method", ["select", "sort"])
def test_small_band_data_bit_exact(small, band_method):
    wl, dem, deltas, dists, _ = small
    note(f"small-201: device band data ({band_method}), 180 sectors")
    with make_engine(wl) as e:
        e.set_band_method(band_method)
        e.compute_bands(0, 180)
        if band_method == "select":
            assert e.bands_sorted_count() == 0
        for s in range(180):
            d, x = e.get_band(s)
            np.testing.assert_array_equal(d, deltas[s], err_msg=f"deltas sector {s}")
            np.testing.assert_array_equal(x.view(np.uint32), dists[s].view(np.uint32),
                                          err_msg=f"distances sector {s}")


@pytest.mark.parametrize("sector", [0, 1, 33, 45, 46, 90, 134, 135, 179])
def test_small_axes_bit_exact(oracle, sector):
    note(f"axes 201 sector {sector}")
    with tvs_amd.Engine(201, 100.0, 6700) as e:
        dist, sight_map, order = e.compute_axes(sector)
    od, osm, oso = oracle.axes_compute(201, sector)
    np.testing.assert_array_equal(dist.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(sight_map, osm)
    np.testing.assert_array_equal(order, oso)


@pytest.mark.parametrize("method", ["plain", "batched"])
@pytest.mark.parametrize("dtype", [np.int16, np.float32])
def test_small_tvs_bit_exact(small, method, dtype):
    wl, dem, _, _, cum = small
    note(f"small-201 TVS {method} {np.dtype(dtype).name}")
    with make_engine(wl) as e:
        e.upload(wl.elevations.astype(dtype))
        e.compute_bands(0, 180)
        e.run(0, 180, method)
        got = e.surfaces()
        got64 = e.surfaces_f64()
    np.testing.assert_array_equal(got.view(np.uint32), cum.view(np.uint32))
    if method == "batched":
        assert np.max(np.abs(got64 - cum) / np.maximum(got64, 1e-30)) < 1e-5


def test_full_int16_range_bit_exact(oracle):
    """Heights over the whole int16 range (negative, -32768, 32767): the
    batched walk's typed int16 loads give float(h) exactly, so int16 and f32
    uploads and the oracle agree bit for bit."""
    w, spacing, los = 101, 100.0, 3300
    rng = np.random.default_rng(11)
    elev = rng.integers(-32768, 32768, size=(w, w), dtype=np.int16)
    elev[::7, ::5] = 32767
    elev[3::11, 2::3] = -32768
    elev[50, :] = rng.integers(-3, 4, size=w, dtype=np.int16)
    elev = elev.reshape(-1)
    note("int16 extremes 101x101: oracle band data + walk")
    dem = oracle.dem_new(w, spacing, los)
    deltas, dists = oracle.band_data_all(dem, range(180))
    cum, _ = oracle.walk_sectors(dem, elev.astype(np.float32), 1.65, deltas, dists,
                                 n_threads=THREADS)
    for dtype in (np.int16, np.float32):
        note(f"int16 extremes: batched, {np.dtype(dtype).name} upload")
        with tvs_amd.Engine(w, spacing, los, 5.0, 1.65, 0) as e:
            e.upload(elev.astype(dtype))
            e.compute_bands(0, 180)
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32),
                                          err_msg=np.dtype(dtype).name)


def test_tile_descriptor_reuse_across_entry_points(small):
    """Back-to-back batched runs reuse the tile descriptor already on the
    device; the ring export and the no-allocation run write the same device
    descriptor buffers, so a batched run after either must copy its own
    again.  Every run here starts from zero surfaces and must equal the
    oracle's 180-sector TVS."""
    wl, _, _, _, cum = small
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        for step in ["run", "run", "rings", "run", "noalloc", "run", "rings", "noalloc", "run"]:
            note(f"descriptor reuse: {step}")
            if step == "rings":
                e.compact_rings(17)
                continue
            e.reset_surfaces()
            if step == "run":
                e.run(0, 180, "batched")
            else:
                e.run_noalloc(0, 180)
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32),
                                          err_msg=step)
        # A different sector range, then the full range again.
        e.reset_surfaces()
        e.run(0, 90, "batched")
        e.run(90, 90, "batched")
        np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32))


@pytest.mark.parametrize("lo,hi", [(0, 1500), (0, 4000), (-1000, 3096), (100, 4196),
                                   (0, 5000)])
def test_half_heights_bit_exact(oracle, lo, hi):
    """int16 DEMs whose heights span at most 4096 also get the f16 form
    h - C; the two-point walk forms h - h0 from it in one mixed FMA in every
    wave whose observers all have an exact C - h0, and converts the int16
    heights in the others.  Ranges: C = 0 (every wave); C = 1952 and 2196
    (the low observers' waves fall back); a negative range around C = 0; a
    span of 4097 (no f16 form).  With the form on and off the TVS equals the
    oracle's bit for bit."""
    w, spacing, los = 201, 90.0, 6000
    elev = (demgen.fractal_dem(w, spacing, 2.4, float(hi - lo), seed=lo + hi).astype(np.int64)
            + lo)
    elev[0, 0], elev[-1, -1] = lo, hi
    elev = elev.astype(np.int16).reshape(-1)
    note(f"half heights [{lo}, {hi}]: oracle band data + walk")
    dem = oracle.dem_new(w, spacing, los)
    deltas, dists = oracle.band_data_all(dem, range(180))
    cum, _ = oracle.walk_sectors(dem, elev.astype(np.float32), 1.65, deltas, dists,
                                 n_threads=THREADS)
    with tvs_amd.Engine(w, spacing, los, 5.0, 1.65, 0) as e:
        e.upload(elev)
        assert e.half_heights() == (True, hi - lo <= 4096)
        e.compute_bands(0, 180)
        for on in (True, False):
            e.set_half_heights(on)
            for shape in [(64, 8, True, False, False, True, 1),
                          (128, 4, True, False, False, True, 1)]:
                note(f"half heights [{lo}, {hi}]: on={on}, shape {shape}")
                e.set_walk_shape(*shape)
                e.reset_surfaces()
                e.run(0, 180, "batched")
                np.testing.assert_array_equal(e.surfaces().view(np.uint32),
                                              cum.view(np.uint32), err_msg=f"{on} {shape}")


@pytest.mark.parametrize("terrain", ["flat", "spikes"])
def test_update_branch_extremes(oracle, terrain):
    """The batched walk skips a band's masked update when no lane of the
    wave sees the point.  Flat terrain: every real step of every band is
    visible (the angle -obs/d - d/12742000 rises with the distance), so no
    update is ever skipped.  A field of 30000 m spikes on every third row
    and column: nearly every band meets a spike within its first steps and
    sees nothing after it, so nearly every update is skipped.  Both against
    the oracle, bit for bit, in the default two-point shape and the
    one-point shape (both directions per lane)."""
    w, spacing, los = 121, 90.0, 3600
    elev = np.zeros((w, w), dtype=np.int16)
    if terrain == "spikes":
        elev[::3, ::3] = 30000
        elev[1::3, 2::3] = 120
    elev = elev.reshape(-1)
    note(f"update branch extremes ({terrain}) 121x121: oracle band data + walk")
    dem = oracle.dem_new(w, spacing, los)
    deltas, dists = oracle.band_data_all(dem, range(180))
    cum, _ = oracle.walk_sectors(dem, elev.astype(np.float32), 1.65, deltas, dists,
                                 n_threads=THREADS)
    for shape in [(64, 8, True, False, False, True, 1), (128, 8, True, False, False, True, 0)]:
        note(f"update branch extremes ({terrain}): batched, shape {shape}")
        with tvs_amd.Engine(w, spacing, los, 5.0, 1.65, 0) as e:
            e.upload(elev)
            e.compute_bands(0, 180)
            e.set_walk_shape(*shape)
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32),
                                          err_msg=str(shape))


@pytest.mark.parametrize("case", ["fraction", "tiny", "integer", "wide"])
def test_fractional_heights_fast_divide(oracle, case):
    """Non-integer f32 heights walk with the four-operation divide when every
    height and observer height is 0 or of magnitude >= 2^-17 (then every
    non-zero height difference is >= 2^-40, inside the divide's proof); a
    DEM holding a smaller non-zero height falls back to the IEEE divide.
    Both give the oracle's surfaces bit for bit, as do integer heights, and
    integer heights spanning 3e6 m (too wide for the short divide's check)."""
    w, spacing, los = 101, 100.0, 3300
    rng = np.random.default_rng(21)
    top = 3_000_000 if case == "wide" else 1200
    elev = (rng.integers(0, top, size=w * w) +
            (rng.random(w * w) if case in ("fraction", "tiny") else 0.0)).astype(np.float32)
    elev[::13] = 0.0
    if case in ("fraction", "tiny"):
        elev[5::17] = np.float32(2.0 ** -17)
    if case == "tiny":
        elev[w * 50 + 50] = np.float32(1e-30)
    note(f"{case} f32 heights 101x101: oracle band data + walk")
    dem = oracle.dem_new(w, spacing, los)
    deltas, dists = oracle.band_data_all(dem, range(180))
    cum, _ = oracle.walk_sectors(dem, elev, 1.65, deltas, dists, n_threads=THREADS)
    with tvs_amd.Engine(w, spacing, los, 5.0, 1.65, 0) as e:
        e.upload(elev)
        assert e.fast_divide() == (case != "tiny")
        e.compute_bands(0, 180)
        for shape in ((128, 16, True, False, False, False, False),
                      (128, 8, True, False, False, True, False)):
            e.set_walk_shape(*shape)
            e.reset_surfaces()
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32),
                                          err_msg=f"{case} {shape}")
        # The short divide: integer heights with an enumerable difference set only.
        assert (e.short_divide()[1] > 0) == (case == "integer")


@pytest.mark.parametrize("w,los", [(120, 3200), (122, 3300), (121, 3300), (123, 3200)])
def test_two_points_per_lane_even_and_odd_widths(oracle, w, los):
    """Two points per lane pairs columns (2c, 2c + 1), shifted one column on
    odd rows of an odd-width DEM (whose left-over column then alternates
    between the last and the first); four points per lane take such pairs
    from rows 2r and 2r + 1 (the last row alone when T is odd).  Even and
    odd widths, each with an even and an odd L (100 m grid: L = 32 or 33),
    give the oracle's surfaces."""
    rng = np.random.default_rng(w)
    elev = rng.integers(0, 900, size=w * w).astype(np.int16)
    note(f"two points per lane {w}x{w}, L={los // 100}: oracle band data + walk")
    dem = oracle.dem_new(w, 100.0, los)
    deltas, dists = oracle.band_data_all(dem, range(180))
    cum, _ = oracle.walk_sectors(dem, elev.astype(np.float32), 1.65, deltas, dists,
                                 n_threads=THREADS)
    with tvs_amd.Engine(w, 100.0, los, 5.0, 1.65, 0) as e:
        e.upload(elev)
        e.compute_bands(0, 180)
        for block, unroll, pers, pair, xcd, two in (
                (64, 4, False, True, 1, 1), (128, 8, False, False, 1, 1), (64, 8, True, False, 1, 1),
                (64, 8, False, True, 2, 1), (64, 4, False, True, 1, 2), (64, 4, False, True, 2, 2),
                (64, 4, False, True, 0, 2)):
            note(f"  shape {block}x{unroll} persistent={pers} pair={pair} xcd={xcd} "
                 f"points per lane {2 * two}")
            e.set_walk_shape(block, unroll, xcd, False, pers, pair, two)
            e.reset_surfaces()
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32),
                                          err_msg=f"{w} shape {block}x{unroll} pair={pair}")


@pytest.mark.parametrize("w", [120, 123])
def test_band_leaving_the_dem_is_the_oracles_panic(oracle, w):
    """A 100 m grid with a 3 km line of sight (L = 30) on 120 or 123 columns:
    some band of the reference's band data leaves the elevation array, where
    kernel.rs:164 panics (the oracle's DEM-index status); the engine refuses
    the same sectors with TVS_ERR_BAND_OUT_OF_DEM, before any walk."""
    dem = oracle.dem_new(w, 100.0, 3000)
    elev = np.zeros(w * w, np.int16)
    bad = []
    for s in range(180):
        deltas, dists = oracle.band_data(dem, s)
        try:
            oracle.walk_sectors(dem, elev.astype(np.float32), 1.65, deltas[None], dists[None])
        except oracle.OracleError as err:
            assert "status 4" in str(err)
            bad.append(s)
    assert bad and len(bad) < 180
    note(f"{w}x{w}, L=30: {len(bad)} sectors leave the DEM")
    with tvs_amd.Engine(w, 100.0, 3000, 5.0, 1.65, 0) as e:
        e.upload(elev)
        for s in range(180):
            if s in bad:
                with pytest.raises(tvs_amd.TvsError) as err:
                    e.compute_bands(s, 1)
                assert err.value.status == _lib.ERR_BAND_OUT_OF_DEM, s
            else:
                e.compute_bands(s, 1)


def test_walk_history(small):
    """tvs_get_walk_history: one walk time per run, oldest first, at most 64,
    the last equal to tvs_get_timing's (events recorded at the same points)."""
    wl, _, _, _, _ = small
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        assert e.walk_history(4).size == 0
        for _ in range(3):
            e.run(0, 180, "batched")
        h = e.walk_history(5)
        assert h.size == 3 and np.all(h > 0)
        assert abs(h[-1] - e.timing().walk_ms) <= 0.1 * h[-1] + 0.01
        for _ in range(70):
            e.run(0, 30, "plain")
        assert e.walk_history(100).size == 64


@pytest.mark.parametrize("pair", [False, True])
def test_short_divide_verified_and_exact(small, pair):
    """The three-operation divide is verified for every sector of this DEM
    and gives the oracle's surfaces; switched off, the split-reciprocal
    divide gives the same bits."""
    wl, dem, _, _, cum = small
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        e.set_walk_shape(128, 8, True, both_directions=pair)
        e.run(0, 180, "batched")
        assert e.short_divide() == (True, 180)
        np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32))
        e.set_short_divide(False)
        e.reset_surfaces()
        e.run(0, 180, "batched")
        assert e.short_divide() == (False, 0)
        np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32))


def test_short_divide_reverified_after_changes(small):
    """A new upload or new band data clears the verification; the next run
    checks again (sector count back to 180) and stays exact."""
    wl, dem, deltas, dists, cum = small
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        e.run(0, 180, "batched")
        assert e.short_divide()[1] == 180
        e.upload(wl.elevations.astype(np.float32))      # integral f32: still eligible
        assert e.short_divide()[1] == 0
        e.set_band(7, deltas[7], dists[7])
        e.reset_surfaces()
        e.run(0, 180, "batched")
        assert e.short_divide()[1] == 180
        np.testing.assert_array_equal(e.surfaces(), cum)
        e.upload(wl.elevations.astype(np.float32) + np.float32(0.25))  # not integral
        e.run(0, 180, "batched")
        assert e.short_divide()[1] == 0


def test_small_tvs_from_host_band_data(small):
    """tvs_set_band (band data from outside) feeds the same walk."""
    wl, dem, deltas, dists, cum = small
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        for s in range(180):
            e.set_band(s, deltas[s], dists[s])
        e.run(0, 180, "batched")
        np.testing.assert_array_equal(e.surfaces(), cum)


@pytest.mark.parametrize("zeros", [1, 2, 9])
def test_leading_zero_distances_bit_exact(small, oracle, zeros):
    """Bands whose first steps sit at distance +0 (3601 x 3601 at 30 m has
    four such sectors): those steps take the IEEE divide one at a time and
    the rest of the band the fast loops, in every launch shape, with the
    oracle's surfaces and ring events."""
    wl, dem, deltas, dists, _ = small
    elev = wl.elevations.astype(np.float32)
    sectors = [3, 47, 95, 150]
    d = deltas[sectors].copy()
    x = dists[sectors].copy()
    x[:, :zeros] = 0.0
    want, rings = oracle.walk_sectors(dem, elev, wl.observer_height, d, x,
                                      rings_per_band=64, rings=True, n_threads=THREADS)
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        for i, s in enumerate(sectors):
            e.set_band(s, d[i], x[i])
        for block, unroll, pipe, pers, pair, two in SHAPES:
            note(f"  zeros {zeros} shape {block}x{unroll} pipe={pipe} pers={pers} pair={pair} "
                 f"two={two}")
            e.set_walk_shape(block, unroll, True, pipe, pers, pair, two)
            e.reset_surfaces()
            for s in sectors:
                e.run(s, 1, "batched")
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), want.view(np.uint32),
                                          err_msg=f"shape {block}x{unroll} pair={pair} two={two}")
        s0 = sectors[1]
        counts, events = e.compact_rings(s0)
        np.testing.assert_array_equal(rings_mod.to_records(e.geometry, counts, events, 64),
                                      rings[1])


@pytest.mark.parametrize("heights", ["f32", "int16", "f32 fraction"])
def test_small_gpu_run_mirror(small, oracle, heights):
    """tvs_run_sector_with_band is GPU::run (gpu.rs:387-468): one sector of
    host band data in, cumulative surfaces out (the batched walk on the
    scratch slot, four-operation divide).  A batched run of the context's
    own sectors afterwards is unaffected by the scratch slot's divide."""
    wl, dem, deltas, dists, cum = small
    upload = {"f32": wl.elevations.astype(np.float32), "int16": wl.elevations,
              "f32 fraction": wl.elevations.astype(np.float32) + np.float32(0.375)}[heights]
    elev = upload.astype(np.float32)
    want = np.zeros(dem.computable_points_count, np.float32)
    with make_engine(wl) as e:
        e.upload(upload)
        for s in (0, 17, 91, 17):
            got = e.run_sector_with_band(deltas[s], dists[s])
            oracle.walk_sectors(dem, elev, wl.observer_height, deltas[s:s + 1],
                                dists[s:s + 1], cumulative=want)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32),
                                          err_msg=f"{heights} sector {s}")
        if heights != "f32 fraction":
            e.compute_bands(0, 180)
            e.reset_surfaces()
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32))


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_small_rings_bit_exact(small, oracle, method):
    wl, dem, deltas, dists, _ = small
    first, count = 40, 4
    note(f"small-201 rings {method}")
    R = oracle.ring_count_per_band(wl.rings_per_km, wl.max_line_of_sight)
    _, want = oracle.walk_sectors(dem, wl.elevations.astype(np.float32), wl.observer_height,
                                  deltas[first:first + count], dists[first:first + count],
                                  rings_per_band=R, rings=True)
    with make_engine(wl, capture=count) as e:
        e.upload(wl.elevations)
        e.compute_bands(first, count)
        e.run(first, count, method)
        for r in range(count):
            np.testing.assert_array_equal(e.rings(r), want[r], err_msg=f"slot {r}")
        assert e.ring_overflow() == 0


@pytest.mark.parametrize("shape", [(512, 8, False, False), (64, 8, True, False),
                                   (64, 4, True, True)])
def test_rings_with_any_walk_shape(small, oracle, shape):
    """Ring capture runs the capture shape whatever shape the context is set
    to (a larger, paired or two-point shape once counted its items wrongly):
    rings and surfaces match the oracle."""
    wl, dem, deltas, dists, _ = small
    first, count = 40, 3
    block, unroll, pair, two = shape
    note(f"small-201 rings with walk shape {shape}")
    R = oracle.ring_count_per_band(wl.rings_per_km, wl.max_line_of_sight)
    want_cum, want = oracle.walk_sectors(dem, wl.elevations.astype(np.float32),
                                         wl.observer_height, deltas[first:first + count],
                                         dists[first:first + count], rings_per_band=R,
                                         rings=True)
    with make_engine(wl, capture=count) as e:
        e.upload(wl.elevations)
        e.compute_bands(first, count)
        e.set_walk_shape(block, unroll, True, both_directions=pair, two_points=two)
        e.run(first, count, "batched")
        for r in range(count):
            np.testing.assert_array_equal(e.rings(r), want[r], err_msg=f"slot {r}")
        np.testing.assert_array_equal(e.surfaces(), want_cum)


# --- 1201 x 1201 -------------------------------------------------------------

@pytest.fixture(scope="module")
def main(oracle):
    wl = demgen.main_1201(seed=0)
    note("main-1201: engine bands for 180 sectors")
    e = make_engine(wl)
    e.upload(wl.elevations)
    e.compute_bands(0, 180)
    note(f"main-1201 bands: {e.timing().bands_ms:.1f} ms, {e.bands_sorted_count()} sorted")
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    yield wl, dem, e
    e.close()


SAMPLE = [0, 1, 2, 44, 45, 46, 89, 90, 91, 134, 135, 136, 178, 179]


def test_main_band_data_bit_exact(main, oracle):
    wl, dem, e = main
    with make_engine(wl) as es:
        es.set_band_method("sort")
        es.compute_bands(0, 180)
        note(f"main-1201 bands by sorts: {es.timing().bands_ms:.1f} ms")
        for s in range(180):       # the two device paths agree everywhere
            a, b = e.get_band(s), es.get_band(s)
            np.testing.assert_array_equal(a[0], b[0], err_msg=f"sector {s}")
            np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    for s in SAMPLE:
        note(f"main-1201 band sector {s}")
        want_d, want_x = oracle.band_data(dem, s)
        d, x = e.get_band(s)
        np.testing.assert_array_equal(d, want_d, err_msg=f"deltas sector {s}")
        np.testing.assert_array_equal(x.view(np.uint32), want_x.view(np.uint32))


@pytest.mark.parametrize("sector", [0, 45, 100])
def test_main_axes_bit_exact(main, oracle, sector):
    wl, dem, e = main
    note(f"main-1201 axes sector {sector}")
    dist, sight_map, order = e.compute_axes(sector)
    od, osm, oso = oracle.axes_compute(wl.width, sector)
    np.testing.assert_array_equal(dist.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(sight_map, osm)
    np.testing.assert_array_equal(order, oso)


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_main_tvs_sample_bit_exact(main, oracle, method):
    wl, dem, e = main
    sectors = [0, 45, 91]
    note(f"main-1201 TVS sample {method}")
    elev = wl.elevations.astype(np.float32)
    want = np.zeros(dem.computable_points_count, np.float32)
    e.reset_surfaces()
    for s in sectors:
        d, x = oracle.band_data(dem, s)
        oracle.walk_sectors(dem, elev, wl.observer_height, d[None], x[None],
                            n_threads=THREADS, cumulative=want)
        e.run(s, 1, method)
    np.testing.assert_array_equal(e.surfaces().view(np.uint32), want.view(np.uint32))


def test_main_full_run_methods_agree(main):
    """All 180 sectors: batched int16 == plain f32 bit for bit, f64 sums
    within 1e-5 of the f32 ones, every surface positive."""
    wl, dem, e = main
    note("main-1201 full plain")
    e.reset_surfaces()
    e.run(0, 180, "plain")
    plain = e.surfaces()
    note("main-1201 full batched")
    e.reset_surfaces()
    e.run(0, 180, "batched")
    batched = e.surfaces()
    b64 = e.surfaces_f64()
    np.testing.assert_array_equal(batched.view(np.uint32), plain.view(np.uint32))
    assert np.all(batched > 0)
    assert np.max(np.abs(b64 - batched) / b64) < 1e-5
    assert e.short_divide() == (True, 180)      # every sector on the short divide
    note("main-1201 full batched, split-reciprocal divide")
    e.set_short_divide(False)
    e.reset_surfaces()
    e.run(0, 180, "batched")
    np.testing.assert_array_equal(e.surfaces().view(np.uint32), plain.view(np.uint32))
    e.set_short_divide(True)


def test_main_descriptor_reuse_after_noalloc_chunks(main):
    """At 1201 x 1201 the no-allocation run walks the 180 sectors in four
    chunks, the last one 24 sectors wide, through the same device descriptor
    buffers as the batched run: a batched run right after it must copy its
    own descriptor again (a stale one would walk the last chunk only).  The
    same after a ring export of one sector.  Compared with a batched run
    from a fresh descriptor state (the plain walk equals it, test above)."""
    wl, dem, e = main
    e.reset_surfaces()
    e.run(0, 180, "batched")
    want = e.surfaces()
    for step in ["noalloc", "run", "rings", "run", "run"]:
        note(f"main-1201 descriptor reuse: {step}")
        if step == "rings":
            e.compact_rings(91)
            continue
        e.reset_surfaces()
        if step == "run":
            e.run(0, 180, "batched")
        else:
            e.run_noalloc(0, 180)
        np.testing.assert_array_equal(e.surfaces().view(np.uint32), want.view(np.uint32),
                                      err_msg=step)


def test_main_rings_one_sector(main, oracle):
    wl, dem, _ = main
    s = 46
    note("main-1201 rings sector 46")
    R = oracle.ring_count_per_band(wl.rings_per_km, wl.max_line_of_sight)
    d, x = oracle.band_data(dem, s)
    _, want = oracle.walk_sectors(dem, wl.elevations.astype(np.float32), wl.observer_height,
                                  d[None], x[None], rings_per_band=R, rings=True)
    with make_engine(wl, capture=1) as e:
        e.upload(wl.elevations)
        e.set_band(s, d, x)
        e.run(s, 1, "batched")
        np.testing.assert_array_equal(e.rings(0), want[0])


def test_main_ring_events_all_sectors(main, oracle):
    """All 180 sectors of main-1201 (321 602 bands each): the compact ring
    export against the oracle's records, band for band (count slot, PoV
    slot, the alternating close/open DEM ids; kernel.rs:141-243).  No band
    fills its R = 180 slots."""
    wl, dem, e = main
    from tvs_amd import rings
    R = oracle.ring_count_per_band(wl.rings_per_km, wl.max_line_of_sight)
    povs = rings.band_povs(e.geometry)
    elev = wl.elevations.astype(np.float32)
    for s in range(180):
        if s % 30 == 0:
            note(f"main-1201 ring events sectors {s}..{s + 29}")
        counts, events = e.compact_rings(s)
        assert int(counts.max()) < R
        d, x = oracle.band_data(dem, s)
        _, want = oracle.walk_sectors(dem, elev, wl.observer_height, d[None], x[None],
                                      rings_per_band=R, rings=True, n_threads=THREADS)
        want = want[0].reshape(-1, R)
        np.testing.assert_array_equal(want[:, 0], counts, err_msg=f"sector {s}")
        np.testing.assert_array_equal(want[:, 1], povs)
        slots = np.arange(2, R)[None, :] <= want[:, :1].astype(np.int64)
        np.testing.assert_array_equal(want[:, 2:][slots], events, err_msg=f"sector {s}")
        assert not np.any(want[:, 2:][~slots])


# --- error paths through the ABI --------------------------------------------

def test_errors_are_status_codes():
    with pytest.raises(tvs_amd.TvsError) as err:
        tvs_amd.Engine(9, 1.0, 5)
    assert err.value.status == _lib.ERR_LOS_TOO_LONG
    with tvs_amd.Engine(9, 1.0, 3) as e:
        with pytest.raises(tvs_amd.TvsError) as err:
            e.run(0, 1)
        assert err.value.status == _lib.ERR_STATE          # no elevations
        with pytest.raises(tvs_amd.TvsError) as err:
            e.upload(np.zeros(80, np.int16))
        assert err.value.status == _lib.ERR_SIZE
        e.upload(np.zeros(81, np.int16))
        with pytest.raises(tvs_amd.TvsError) as err:
            e.run(0, 1)
        assert err.value.status == _lib.ERR_STATE          # no band data
        with pytest.raises(tvs_amd.TvsError) as err:
            e.run(170, 20)
        assert err.value.status == _lib.ERR_INVALID_ARGUMENT
        with pytest.raises(tvs_amd.TvsError) as err:
            e.set_band(0, np.array([1000, 0, 0]), np.array([1.0, 0, 0], np.float32))
        assert err.value.status == _lib.ERR_BAND_OUT_OF_DEM
    with tvs_amd.Engine(9, 1.0, 1) as e:                   # band_size 2
        with pytest.raises(tvs_amd.TvsError) as err:
            e.compute_bands(0, 1)
        assert err.value.status == _lib.ERR_BAND_TOO_SHORT


def test_ring_overflow_is_counted_not_spilled(small):
    """With too few reserved rings the reference writes into the next band
    (kernel.rs:141); the engine drops and counts instead, and the surfaces
    are unaffected."""
    wl, dem, _, _, cum = small
    with make_engine(wl, capture=180, rings_per_km=0.5) as e:
        assert e.rings_per_band == 3
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        e.run(0, 180, "batched")
        assert e.ring_overflow() > 0
        np.testing.assert_array_equal(e.surfaces(), cum)
        r = e.rings(0).reshape(-1, 3)
        assert np.all(r[:, 0] % 2 == 0)


SHAPES = [  # (block, unroll, pipeline, persistent, both directions, two points per lane)
    (256, 8, False, False, False, False), (128, 8, False, False, False, False),
    (64, 8, False, False, False, False), (512, 8, False, False, False, False),
    (256, 4, False, False, False, False), (256, 16, False, False, False, False),
    (128, 16, False, False, False, False), (256, 8, True, False, False, False),
    (128, 8, True, False, False, False), (256, 4, True, False, False, False),
    (512, 8, True, False, False, False), (256, 8, False, True, False, False),
    (128, 8, False, True, False, False), (64, 8, False, True, False, False),
    (64, 8, False, False, True, False), (128, 8, False, False, True, False),
    (256, 8, False, False, True, False), (64, 4, False, False, True, False),
    (128, 4, False, False, True, False), (64, 8, False, False, False, True),
    (128, 8, False, False, False, True), (64, 4, False, False, False, True),
    (64, 4, False, False, True, True), (128, 4, False, False, True, True),
    (64, 8, False, False, True, True), (64, 4, False, True, True, True),
    (64, 8, False, True, False, True), (64, 4, False, False, True, 2)]


# xcd: 0 round-robin items, 1 XCD-contiguous, 2 XCD-contiguous in the
# directional order (each sector's blocks sorted across its band direction).
@pytest.mark.parametrize("xcd", [1, 0, 2])
def test_every_walk_shape_bit_exact(small, xcd):
    """Each compiled launch shape of the batched walk, in each item order,
    gives the oracle's surfaces (the tuner may pick any of them)."""
    wl, dem, _, _, cum = small
    with make_engine(wl) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        for block, unroll, pipe, pers, pair, two in SHAPES:
            note(f"shape {block}x{unroll} pipe={pipe} pers={pers} pair={pair} two={two} "
                 f"xcd={xcd}")
            e.set_walk_shape(block, unroll, xcd, pipe, pers, pair, two)
            e.reset_surfaces()
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces(), cum,
                                          err_msg=f"shape {block}x{unroll} pipe={pipe} "
                                                  f"pers={pers} pair={pair} two={two}")


@pytest.fixture(scope="module", params=["spaced", "tiny"])
def small_fractional(request, oracle):
    """small-201 seed 0 plus a fractional part per point: f32 heights that
    stay f32 on the device (integer-valued f32 uploads are stored as int16).
    "tiny" also holds one height of 1e-30, which sends the DEM to the IEEE
    divide."""
    wl = demgen.small_201(seed=0)
    rng = np.random.default_rng(5)
    elev = (wl.elevations.astype(np.float64) + 0.125 + 0.75 * rng.random(wl.elevations.size)
            ).astype(np.float32)
    if request.param == "tiny":
        elev[wl.width * 100 + 37] = np.float32(1e-30)
    note(f"small-201 fractional ({request.param}): oracle band data + walk")
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    deltas, dists = oracle.band_data_all(dem, range(180))
    cum, _ = oracle.walk_sectors(dem, elev, wl.observer_height, deltas, dists,
                                 n_threads=THREADS)
    return wl, elev, cum, request.param == "spaced"


@pytest.mark.parametrize("xcd", [1, 0, 2])
def test_every_walk_shape_bit_exact_f32(small_fractional, xcd):
    """The f32-height kernels of every compiled shape, on non-integer
    heights (four-operation divide, or IEEE with a tiny height), give the
    oracle's surfaces."""
    wl, elev, cum, fast = small_fractional
    with make_engine(wl) as e:
        e.upload(elev)
        assert e.fast_divide() == fast
        e.compute_bands(0, 180)
        for block, unroll, pipe, pers, pair, two in SHAPES:
            note(f"shape {block}x{unroll} pipe={pipe} pers={pers} pair={pair} two={two} "
                 f"xcd={xcd}")
            e.set_walk_shape(block, unroll, xcd, pipe, pers, pair, two)
            e.reset_surfaces()
            e.run(0, 180, "batched")
            np.testing.assert_array_equal(e.surfaces(), cum,
                                          err_msg=f"shape {block}x{unroll} pipe={pipe} "
                                                  f"pers={pers} pair={pair} two={two}")


def test_directional_order_shape_round_trip(small):
    """xcd_contiguous 2 selects the directional item order and reads back as
    such; other values above 1 are refused."""
    wl = small[0]
    with make_engine(wl) as e:
        e.set_walk_shape(64, 8, 2, False, False, True, True)
        shape = e.walk_shape()
        assert shape["xcd_contiguous"] and shape["directional"]
        e.set_walk_shape(64, 8, 1, False, False, True, True)
        assert not e.walk_shape()["directional"]
        with pytest.raises(tvs_amd.TvsError) as err:
            e.set_walk_shape(64, 8, 3, False, False, True, True)
        assert err.value.status == _lib.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("w,heights,directional", [(2896, "int16", False),
                                                   (2897, "int16", True),
                                                   (2048, "f32 fraction", False),
                                                   (2049, "f32 fraction", True)])
def test_default_item_order_by_dem_size(w, heights, directional):
    """Untuned DEMs whose heights take more than 16 MiB (int16 above 2^23
    points: 2897^2, not 2896^2; f32 above 2^22: 2049^2, not 2048^2) walk in
    the directional item order by default, smaller ones in the
    XCD-contiguous order; a set shape overrides it."""
    elev = np.random.default_rng(w).integers(0, 2000, w * w).astype(np.int16)
    if heights != "int16":
        elev = elev.astype(np.float32) + np.float32(0.25)
    with tvs_amd.Engine(w, 30.0, (w // 3) * 30, 5.0, 1.65, 0) as e:
        e.upload(elev)
        shape = e.walk_shape()
        note(f"default shape {w}x{w} {heights}: {shape}")
        assert shape["xcd_contiguous"] and shape["directional"] == directional
        e.set_walk_shape(64, 8, 1, False, False, True, 1)
        assert not e.walk_shape()["directional"]
