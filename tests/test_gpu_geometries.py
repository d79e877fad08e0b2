"""Seeded random geometries (64 by default) through the ABI against the oracle, bit for bit:
widths odd and even, L of either parity, four grid spacings, observer
heights from 0 to 30 m, terrain from flat-ish to the whole int16 range,
int16, integer-valued f32 and non-integer f32 uploads; the plain walk, the
batched walk in the default and a drawn launch shape, and one sector's
ring events.  Sectors whose band leaves the DEM (the reference's
kernel.rs:164 panic) must be refused, as the oracle refuses them."""
import os
import sys

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen, rings, _lib

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_parity import SHAPES  # noqa: E402

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def draw(seed: int):
    rng = np.random.default_rng(1000 + seed)
    w = int(rng.integers(31, 171))
    spacing = float(rng.choice([10.0, 30.0, 90.0, 100.0]))
    L = int(rng.integers(3, max(4, w // 3) + 1))
    los = int(L * spacing)
    obs = float(rng.choice([0.0, 1.65, 1.8, 30.0]))
    kind = ["fractal", "noise", "extreme"][seed % 3]
    if kind == "fractal":
        elev = demgen.fractal_dem(w, spacing, 2.4, float(rng.integers(10, 4000)), seed)
    elif kind == "noise":
        elev = rng.integers(-200, 2500, size=(w, w)).astype(np.int16)
    else:
        elev = rng.integers(-32768, 32768, size=(w, w)).astype(np.int16)
    elev = elev.reshape(-1).astype(np.int16)
    upload = ["int16", "f32", "f32 fraction"][(seed // 3) % 3]
    if upload == "int16":
        heights = elev
    elif upload == "f32":
        heights = elev.astype(np.float32)
    else:
        heights = (elev.astype(np.float64) + rng.random(elev.size) * 0.9 + 0.05).astype(np.float32)
    shape = SHAPES[int(rng.integers(len(SHAPES)))]
    return w, spacing, los, obs, kind, upload, heights, shape


# TVS_GEOMETRY_SEEDS sets the number of geometries (64 by default).
@pytest.mark.parametrize("seed", range(int(os.environ.get("TVS_GEOMETRY_SEEDS", "64"))))
def test_random_geometry_bit_exact(oracle, seed):
    w, spacing, los, obs, kind, upload, heights, shape = draw(seed)
    print(f"seed {seed}: {w}x{w} @ {spacing} m, los {los} m, obs {obs} m, {kind}, "
          f"{upload}, shape {shape}", flush=True)
    dem = oracle.dem_new(w, spacing, los)
    deltas, dists = oracle.band_data_all(dem, range(180))
    f32 = heights.astype(np.float32)
    ok = []
    cum = np.zeros(dem.computable_points_count, np.float32)
    for s in range(180):
        tmp = cum.copy()                           # a refused sector adds nothing
        try:
            oracle.walk_sectors(dem, f32, obs, deltas[s:s + 1], dists[s:s + 1],
                                n_threads=THREADS, cumulative=tmp)
        except oracle.OracleError as err:
            assert "status 4" in str(err)          # the band leaves the DEM
            continue
        cum = tmp
        ok.append(s)
    # Every sector the oracle walks, in order, accumulated in `cum`.
    with tvs_amd.Engine(w, spacing, los, 5.0, obs, 0) as e:
        e.upload(heights)
        for s in range(180):
            if s in ok:
                e.compute_bands(s, 1)
            else:
                with pytest.raises(tvs_amd.TvsError) as err:
                    e.compute_bands(s, 1)
                assert err.value.status == _lib.ERR_BAND_OUT_OF_DEM, s
        if not ok:
            return
        # The first walkable sector's ring events against the oracle's
        # records (64 slots a band, compact export rebuilt to that layout).
        s0 = ok[0]
        counts, events = e.compact_rings(s0)
        _, want = oracle.walk_sectors(dem, f32, obs, deltas[s0:s0 + 1], dists[s0:s0 + 1],
                                      rings_per_band=64, rings=True, n_threads=THREADS)
        np.testing.assert_array_equal(rings.to_records(e.geometry, counts, events, 64), want[0],
                                      err_msg=f"seed {seed} rings sector {s0}")
        for shp in ("plain", None, shape):
            method = "plain" if shp == "plain" else "batched"
            if shp not in (None, "plain"):
                block, unroll, pipe, pers, pair, two = shp
                e.set_walk_shape(block, unroll, True, pipe, pers, pair, two)
            e.reset_surfaces()
            for a in range(len(ok)):               # contiguous runs of walkable sectors
                if a == 0 or ok[a] != ok[a - 1] + 1:
                    b = a
                    while b + 1 < len(ok) and ok[b + 1] == ok[b] + 1:
                        b += 1
                    e.run(ok[a], b - a + 1, method)
            np.testing.assert_array_equal(e.surfaces().view(np.uint32), cum.view(np.uint32),
                                          err_msg=f"seed {seed} shape {shp}")
