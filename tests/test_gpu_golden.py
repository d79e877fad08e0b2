"""The reference's own test vectors, reproduced by the HIP engine through the
C ABI (axes.rs, band_of_sight.rs, cache.rs and compute.rs tests)."""
import sys

import numpy as np
import pytest

import tvs_amd

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", [0, 1])
def test_axes_golden(golden, case):
    c = golden["axes"][case]
    note(f"axes golden {c['source']}")
    with tvs_amd.Engine(c["width"], 1.0, 1) as e:
        dist, sight_map, sector_ordered = e.compute_axes(c["sector"])
    np.testing.assert_array_equal(dist, np.array(c["distances"], np.float32))
    order = np.empty_like(sight_map)
    order[sight_map] = np.arange(sight_map.size, dtype=np.uint32)
    assert order.tolist() == c["sight_order"]
    assert sector_ordered.tolist() == c["sector_ordered"]


def reconstruct(dem, deltas):
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
def test_band_golden(golden, case):
    c = golden["bands"]["cases"][case]
    note(f"band golden {c['source']}")
    dem = tvs_amd.DEM(9, 1.0, 3)
    with tvs_amd.Engine(9, 1.0, 3) as e:
        e.compute_bands(c["sector"], 1)
        deltas, dists = e.get_band(c["sector"])
    assert reconstruct(dem, deltas) == c["bands"]
    np.testing.assert_array_equal(dists, np.array(c["distances"], np.float32))


def test_cache_35_golden(golden):
    c = golden["cache_35"]
    with tvs_amd.Engine(9, 1.0, 3) as e:
        e.compute_bands(c["sector"], 1)
        deltas, _ = e.get_band(c["sector"])
        dist, _, _ = e.compute_axes(c["sector"])
    assert deltas.tolist() == c["band_deltas"]
    assert dist[40] == np.float32(c["axes_distance_40"])
    assert dist[-1] == np.float32(c["axes_distance_last"])


@pytest.mark.parametrize("method", ["hip", "hip-plain"])
@pytest.mark.parametrize("name", ["single_peak", "double_peak"])
@pytest.mark.parametrize("dtype", [np.int16, np.float32])
def test_tvs_golden(golden, method, name, dtype):
    t = golden["tvs"]
    note(f"tvs golden {name} {method} {np.dtype(dtype).name}")
    dem = tvs_amd.DEM(9, 1.0, 3)
    dem.elevations = np.array(t[name + "_dem"]["elevations"], dtype)
    compute = tvs_amd.Compute(method, None, None, dem, 5000.0, observer_height=1.8)
    surfaces, _ = compute.run()
    np.testing.assert_array_equal(surfaces, np.array(t[name + "_totals"]["values"], np.float32))
    assert compute.engine.ring_overflow() == 0


@pytest.mark.parametrize("method", ["hip", "hip-plain"])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_viewshed_golden(golden, oracle, method, case):
    v = golden["viewsheds"][case]
    t = golden["tvs"]
    note(f"viewshed golden {v['source']} {method}")
    dem = tvs_amd.DEM(9, 1.0, 3)
    dem.elevations = np.array(t[v["dem"]]["elevations"], np.int16)
    compute = tvs_amd.Compute(method, None, None, dem, 5000.0, observer_height=1.8)
    _, rings = compute.run()
    odem = oracle.dem_new(9, 1.0, 3)
    rows = oracle.output_ascii(odem, v["pov_id"], rings, compute.reserved_rings_per_band)
    assert rows == v["rows"]
