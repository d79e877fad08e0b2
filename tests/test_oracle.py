"""The oracle (CPU restatement of the reference path) against the reference's
own golden vectors, brute force and closed forms.  CPU only."""
import itertools
import struct

import numpy as np
import pytest


def total_order(f: float) -> int:
    u = struct.unpack("<I", struct.pack("<f", f))[0]
    return (~u & 0xFFFFFFFF) if u & 0x80000000 else (u | 0x80000000)


# --- axes.rs tests -------------------------------------------------------

@pytest.mark.parametrize("case", [0, 1])
def test_axes_golden(oracle, golden, case):
    c = golden["axes"][case]
    dist, sight_map, sector_ordered = oracle.axes_compute(c["width"], c["sector"])
    np.testing.assert_array_equal(dist, np.array(c["distances"], np.float32))
    order = np.empty_like(sight_map)
    order[sight_map] = np.arange(sight_map.size, dtype=np.uint32)  # extract_from_mapped
    assert order.tolist() == c["sight_order"]
    assert sector_ordered.tolist() == c["sector_ordered"]


def test_order_by_distance_is_stable_total_order(oracle):
    """Brute force: Python's stable sort on the IEEE total-order key of the
    f32-rounded distance, with ties (including -0/+0) kept in id order."""
    rng = np.random.default_rng(7)
    for n in (1, 2, 17, 300, 4099):
        d = rng.integers(-6, 6, n).astype(np.float64) * 0.25
        d[rng.random(n) < 0.1] = -0.0
        d[rng.random(n) < 0.1] = 0.0
        d += rng.random(n) * 1e-12          # lost in the f32 rounding: ties
        got = oracle.order_by_distance(d)
        want = sorted(range(n), key=lambda i: (total_order(float(np.float32(d[i]))), i))
        assert got.tolist() == want


def test_negative_zero_sorts_before_positive_zero(oracle):
    got = oracle.order_by_distance(np.array([0.0, -0.0, 0.0, -0.0]))
    assert got.tolist() == [1, 3, 0, 2]


def test_rotated_distance_formula_fixed_points(oracle):
    """axes.rs:154-176 at the shifted 0 degrees: row r, column x gives
    x*sin(a) + r*cos(a); at 90.001 degrees the roles swap."""
    w = 6
    a = np.float32(0.001)
    d = oracle.calculate_distances(w, float(a)).reshape(w, w)
    rad = float(a) * (np.pi / 180.0)
    for r, x in itertools.product(range(w), range(w)):
        assert d[r, x] == x * np.sin(rad) - (-r) * np.cos(rad)


# --- band_of_sight.rs / cache.rs tests -----------------------------------

def reconstruct(dem, deltas):
    """band_of_sight.rs:198-228"""
    bands = []
    for p in range(dem.computable_points_count):
        f = b = dem.tvs_id_to_pov_id(p)
        front, back = [f], [b]
        for x in deltas:
            f += int(x)
            b -= int(x)
            front.append(f)
            back.append(b)
        bands.append(back[::-1])
        bands.append(front)
    return bands


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_band_golden(oracle, golden, case):
    c = golden["bands"]["cases"][case]
    dem = oracle.dem_new(9, 1.0, 3)
    assert dem.computable_points_count == 9
    deltas, dists = oracle.band_data(dem, c["sector"])
    assert reconstruct(dem, deltas) == c["bands"]
    np.testing.assert_array_equal(dists, np.array(c["distances"], np.float32))


def test_cache_35_golden(oracle, golden):
    c = golden["cache_35"]
    dem = oracle.dem_new(9, 1.0, 3)
    dist, sight_map, sector_ordered = oracle.axes_compute(9, c["sector"])
    deltas, _ = oracle.compile_band_data(dem, dist, sight_map, sector_ordered)
    assert deltas.tolist() == c["band_deltas"]
    assert dist.size == c["axes_distances_len"]
    assert dist[40] == np.float32(c["axes_distance_40"])
    assert dist[-1] == np.float32(c["axes_distance_last"])


# --- compute.rs tests ----------------------------------------------------

@pytest.mark.parametrize("name", ["single_peak", "double_peak"])
def test_tvs_golden(oracle, golden, name):
    t = golden["tvs"]
    dem = oracle.dem_new(t["dem_width"], t["scale"], t["max_line_of_sight"])
    elev = np.array(t[name + "_dem"]["elevations"], np.float32)
    cum, _, spill = oracle.compute_run(dem, elev, t["observer_height"], t["rings_per_km"])
    np.testing.assert_array_equal(cum, np.array(t[name + "_totals"]["values"], np.float32))
    assert spill == 0


def test_tvs_golden_detects_a_planted_change(oracle, golden):
    """The golden totals are sensitive to the observer height (compute.rs:41
    hardcodes 1.8; the CLI default is 1.65): the same oracle at 1.65 must
    miss them, so a dropped or altered observer term cannot pass."""
    t = golden["tvs"]
    dem = oracle.dem_new(9, 1.0, 3)
    elev = np.array(t["single_peak_dem"]["elevations"], np.float32)
    cum, _, _ = oracle.compute_run(dem, elev, 1.65, 5000.0, rings=False)
    assert not np.array_equal(cum, np.array(t["single_peak_totals"]["values"], np.float32))


@pytest.mark.parametrize("case", [0, 1, 2])
def test_viewshed_golden(oracle, golden, case):
    v = golden["viewsheds"][case]
    t = golden["tvs"]
    dem = oracle.dem_new(9, 1.0, 3)
    elev = np.array(t[v["dem"]]["elevations"], np.float32)
    _, ring, _ = oracle.compute_run(dem, elev, 1.8, 5000.0)
    R = oracle.ring_count_per_band(5000.0, 3)
    assert oracle.output_ascii(dem, v["pov_id"], ring, R) == v["rows"]


# --- geometry (dem.rs, band_of_sight.rs) ---------------------------------

@pytest.mark.parametrize("w,scale,los,L,T,N,steps,real,R", [
    (9, 1.0, 3, 3, 3, 9, 3, 3, 0),
    (201, 100.0, 6700, 67, 67, 4489, 67, 40, 33),
    (1201, 90.0, 36000, 400, 401, 160801, 400, 235, 180),
    (3601, 30.0, 40000, 1333, 933, 870489, 1333, 782, 200),
    (257, 30.0, 2550, 85, 87, 7569, 85, 51, 12),
    (257, 90.0, 7650, 85, 87, 7569, 85, 51, 38),
])
def test_config_geometry(oracle, w, scale, los, L, T, N, steps, real, R):
    dem = oracle.dem_new(w, scale, los)
    assert (dem.max_los_as_points, dem.tvs_width, dem.computable_points_count) == (L, T, N)
    assert dem.n_deltas == steps
    if R:
        assert oracle.ring_count_per_band(5.0, los) == R
    if w <= 1201:
        deltas, dists = oracle.band_data(dem, 0)
        # Only the first 2B - bs/2 steps are filled (band_of_sight.rs:149-169)
        bs = int(2 * (L + 1) * np.sqrt(2.0))
        assert 2 * (L + 1) - bs // 2 == real
        assert np.count_nonzero(deltas) == real
        assert np.all(deltas[real:] == 0) and np.all(dists[real:] == 0)


def test_los_too_long_is_an_error(oracle):
    with pytest.raises(oracle.OracleError):
        oracle.dem_new(9, 1.0, 5)


def test_short_band_panics_become_errors(oracle):
    """band_size <= 2 indexes past band_deltas (band_of_sight.rs:165)."""
    for los in (0, 1):
        dem = oracle.dem_new(9, 1.0, los)
        with pytest.raises(oracle.OracleError):
            oracle.band_data(dem, 0)


# --- invariants ----------------------------------------------------------

def test_flat_dem_closed_form(oracle):
    """On a flat DEM the walk does not depend on the observer and forward
    equals backward, so every TVS is the same f32 number: the in-order sum
    over sectors of (s + s) where s is the band surface from the band data
    alone.  Each real step with a larger distance than all before it is
    visible (the angle -h/d - d/12.742e6 grows with d far below
    sqrt(h * 12.742e6) grid steps)."""
    w, los = 61, 20
    dem = oracle.dem_new(w, 1.0, los)
    elev = np.full(w * w, 250.0, np.float32)
    sectors = range(0, 180, 7)
    cum = np.zeros(dem.computable_points_count, np.float32)
    want = np.float32(0.0)
    for s in sectors:
        deltas, dists = oracle.band_data(dem, s)
        h0 = np.float32(250.0) + np.float32(1.65)
        mx = np.float32(-2000.0)
        surf = np.float32(0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            for d in dists:
                ang = np.float32(np.float32(np.float32(250.0) - h0) / d) - np.float32(d / np.float32(12742000.0))
                if ang > mx:
                    surf = np.float32(surf + np.float32(d * np.float32(0.0174533)))
                mx = np.float32(np.fmax(ang, mx))
        want = np.float32(np.float32(want + surf) + surf)
        oracle.compute_run(dem, elev, 1.65, 5.0, first_sector=s, sector_count=1,
                           rings=False, cumulative=cum)
    assert np.all(cum == want)


def test_ring_counts_are_even(oracle):
    """Every opening, including the implicit one at the observer, is closed
    (kernel.rs:145, 220-243): ring[kR] is even for every band."""
    rng = np.random.default_rng(3)
    dem = oracle.dem_new(31, 1.0, 10)
    elev = rng.integers(0, 40, 31 * 31).astype(np.float32)
    _, ring, spill = oracle.compute_run(dem, elev, 1.65, 1000.0, sector_count=12)
    R = oracle.ring_count_per_band(1000.0, 10)
    assert spill == 0
    counts = ring.reshape(12, -1, R)[:, :, 0]
    assert np.all(counts % 2 == 0) and np.all(counts >= 2)


def test_threaded_walk_matches_serial(oracle):
    rng = np.random.default_rng(5)
    dem = oracle.dem_new(41, 10.0, 130)
    elev = rng.integers(0, 300, 41 * 41).astype(np.float32)
    a, _, _ = oracle.compute_run(dem, elev, 1.65, 5.0, sector_count=30, rings=False, n_threads=1)
    b, _, _ = oracle.compute_run(dem, elev, 1.65, 5.0, sector_count=30, rings=False, n_threads=4)
    np.testing.assert_array_equal(a, b)


def test_concurrent_band_data_matches_serial(oracle):
    """Band data compiled for many sectors from several threads at once (as
    the bench's all-cores baseline does) equals the serial result."""
    from concurrent.futures import ThreadPoolExecutor
    dem = oracle.dem_new(101, 30.0, 1000)
    d1, x1 = oracle.band_data_all(dem, range(180))
    with ThreadPoolExecutor(8) as ex:
        parts = list(ex.map(lambda s: oracle.band_data_all(dem, [s]), range(180)))
    np.testing.assert_array_equal(np.concatenate([p[0] for p in parts]), d1)
    np.testing.assert_array_equal(np.concatenate([p[1] for p in parts]).view(np.uint32),
                                  x1.view(np.uint32))


def test_f64_companion(oracle):
    """The f64 walk agrees with the f32 walk to well within 1e-4 and sees
    no visibility flips on a flat DEM."""
    rng = np.random.default_rng(11)
    dem = oracle.dem_new(61, 30.0, 600)
    elev = rng.integers(0, 500, 61 * 61).astype(np.float32)
    deltas, dists = oracle.band_data_all(dem, range(0, 180, 9))
    cum, _ = oracle.walk_sectors(dem, elev, 1.65, deltas, dists, n_threads=4)
    tvs64, flips, steps = oracle.walk_sectors_f64(dem, elev, 1.65, deltas, dists, n_threads=4)
    assert steps == 20 * 2 * dem.computable_points_count * dem.n_deltas
    assert np.max(np.abs(cum - tvs64) / tvs64) < 1e-4
    flat = np.full(61 * 61, 100.0, np.float32)
    _, flips_flat, _ = oracle.walk_sectors_f64(dem, flat, 1.65, deltas, dists)
    assert flips_flat == 0


def test_f64_flat_dem_closed_form(oracle):
    """On a flat DEM the f64 angle -h/d - d/12.742e6 (h = the observer
    height, exact in double) grows strictly with d for every d far below
    sqrt(h * 12.742e6) ~ 4585 grid steps, and distinct f32 distances differ
    by far more than double rounding can hide.  So a step is visible in f64
    exactly when its distance exceeds every distance before it in the band
    (a zero distance gives -inf and is never visible), and each TVS value is
    the in-order double sum over sectors of (s + s), s the sum of
    d * 0.0174533 over those record-breaking steps -- from band data alone,
    with no angle, division or max-scan computed."""
    w, los = 61, 20
    dem = oracle.dem_new(w, 1.0, los)
    elev = np.full(w * w, 250.0, np.float32)
    sectors = list(range(0, 180, 7))
    deltas, dists = oracle.band_data_all(dem, sectors)
    want = 0.0
    for row in dists:
        surf, longest = 0.0, 0.0
        for d in row.astype(np.float64):
            if d > longest:
                surf += d * 0.0174533
            longest = max(longest, d)
        want += surf
        want += surf
    tvs64, flips, _ = oracle.walk_sectors_f64(dem, elev, 1.65, deltas, dists, n_threads=4)
    assert flips == 0
    assert np.all(tvs64 == want), (tvs64[:3], want)


def test_f64_constructed_visibility_flip(oracle):
    """A band built so that f32 and f64 provably see differently.

    9x9 DEM (L = 3, observers at x, y in 3..5), one sector of hand-made band
    data: deltas (1, 1, 1) (forward = east along the row), distances
    (2, 1, 3), observer height 0, heights 0 except row 3: x = 4 is 2000 m,
    x = 5 is 1000 m.  With E = 12742000:

    * observer (3, 3), forward: step 1 sees 2000/2 - 2/E; step 2 sees
      1000/1 - 1/E, larger by 1/E = 7.8e-8 in f64 (visible) but rounded to
      the same f32 1000.0 as step 1 (ulp 6.1e-5), and `>` is strict, so
      f32 hides it: the only flip.  Step 3 (0 m at d 3) is hidden.
    * observer (4, 3) (h0 2000): forward -500, -2000, -666.7: step 1 only;
      backward -1000, -2000, -666.7: steps 1 and 3.
    * observer (5, 3) (h0 1000): forward -500, -1000, -333.3: steps 1, 3;
      backward 500, -1000, -333.3: step 1.
    * every other band runs over zero heights from h0 = 0: -2/E, -1/E,
      -3/E, steps 1 and 2 visible in both precisions (f32 keeps full
      relative precision there).

    So flips = 1 of 54 steps, and the f64 surfaces are the double sums of
    d * 0.0174533 over those steps (forward then backward); the f32 walk
    gives the same sets except observer (3, 3) forward, which loses step 2.
    A f64 walk taking its visibility from the f32 angle fails both."""
    w = 9
    dem = oracle.dem_new(w, 1.0, 3)
    assert (dem.max_los_as_points, dem.tvs_width, dem.n_deltas) == (3, 3, 3)
    elev = np.zeros(w * w, np.float32)
    elev[3 * w + 4] = 2000.0
    elev[3 * w + 5] = 1000.0
    deltas = np.array([[1, 1, 1]], np.int32)
    dists = np.array([[2.0, 1.0, 3.0]], np.float32)
    visible = {   # observer t -> (forward steps, backward steps), f64
        0: ((0, 1), (0, 1)), 1: ((0,), (0, 2)), 2: ((0, 2), (0,)),
    }
    want64 = np.zeros(dem.computable_points_count)
    want32 = np.zeros(dem.computable_points_count, np.float32)
    for t in range(dem.computable_points_count):
        fwd, bwd = visible.get(t, ((0, 1), (0, 1)))
        for band in (fwd, bwd):
            s = 0.0
            for i in band:
                s += float(dists[0, i]) * 0.0174533
            want64[t] += s
        for b, band in enumerate((fwd, bwd)):
            f32_band = (0,) if (t == 0 and b == 0) else band
            s32 = np.float32(0.0)
            for i in f32_band:
                s32 = np.float32(s32 + np.float32(dists[0, i] * np.float32(0.0174533)))
            want32[t] = np.float32(want32[t] + s32)
    tvs64, flips, steps = oracle.walk_sectors_f64(dem, elev, 0.0, deltas, dists)
    assert steps == 54
    assert flips == 1
    assert np.array_equal(tvs64, want64), (tvs64, want64)
    cum, _ = oracle.walk_sectors(dem, elev, 0.0, deltas, dists)
    assert np.array_equal(cum, want32), (cum, want32)
