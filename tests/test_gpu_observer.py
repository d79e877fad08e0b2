"""tvs_observer_walk: one observer's bands, one wavefront per band, through
the C ABI, bit for bit: the reference's ASCII viewsheds (compute.rs:357-416
via output/ascii.rs), each band's ring record and surface against the
oracle's kernel (kernel.rs:71-256), the same events as the full-sector
export (tvs_compact_rings), and the 180 band pairs summed in sector order
equal to the observer's TVS from the batched pass."""
import os

import numpy as np
import pytest

import tvs_amd
from tvs_amd import _lib, demgen, rings

pytestmark = pytest.mark.gpu
THREADS = min(16, len(os.sched_getaffinity(0)))


def tvs_id_of(g, pov_id: int) -> int:
    """dem.rs:118-124"""
    x = pov_id % g.dem_width - g.max_los_as_points
    y = pov_id // g.dem_width - g.max_los_as_points
    return y * g.tvs_width + x


def pov_of(g, t: int) -> int:
    """dem.rs:127-132"""
    return (t // g.tvs_width + g.max_los_as_points) * g.dem_width + \
        (t % g.tvs_width + g.max_los_as_points)


@pytest.mark.parametrize("case", [0, 1, 2])
def test_observer_walk_golden_viewsheds(golden, case):
    v = golden["viewsheds"][case]
    t = golden["tvs"]
    print(f"observer walk golden {v['source']}", flush=True)
    with tvs_amd.Engine(9, 1.0, 3, 5000.0, 1.8) as e:
        e.upload(np.array(t[v["dem"]]["elevations"], np.int16))
        e.compute_bands(0, 180)
        counts, _, events = e.observer_walk(tvs_id_of(e.geometry, v["pov_id"]))
        pairs = rings.observer_rings(v["pov_id"], counts, events)
        assert rings.ascii_viewshed(e.geometry, v["pov_id"], pairs=pairs) == v["rows"]


def _sum_in_sector_order(surf):
    """cumulative += forward; cumulative += backward, sector after sector
    (compute.rs:229 with kernel.rs:254), in f32."""
    c = np.float32(0.0)
    for x in surf:
        c = np.float32(c + x)
    return c


@pytest.mark.parametrize("seed", [4, 9])
def test_observer_walk_against_oracle_bands(oracle, seed):
    """small-201: for observers at the corners and inside, every band's ring
    record (count, PoV, events) and surface equal the oracle's kernel for
    that kernel id, in all 180 sectors; the events equal the full-sector
    export's for the same bands."""
    wl = demgen.small_201(seed=seed)
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    deltas, dists = oracle.band_data_all(dem, range(180))
    elev = wl.elevations.astype(np.float32)
    N = dem.computable_points_count
    rng = np.random.default_rng(seed)
    observers = [0, N - 1, dem.tvs_width - 1, int(rng.integers(N)), int(rng.integers(N))]
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        compact = {s: e.compact_rings(s) for s in (0, 37, 90, 179)}
        for t in observers:
            print(f"small-201 seed {seed}: observer {t}", flush=True)
            counts, surf, events = e.observer_walk(t)
            off = rings.offsets(counts)
            pov = pov_of(e.geometry, t)
            for s in range(180):
                for direction in (0, 1):
                    b = 2 * s + direction
                    want_surf, want_rec = oracle.kernel_band(dem, elev, wl.observer_height,
                                                             deltas[s], dists[s], t + direction * N)
                    got_rec = np.concatenate([[counts[b], pov],
                                              events[off[b]:off[b] + counts[b] - 1]]).astype(np.uint32)
                    np.testing.assert_array_equal(got_rec, want_rec,
                                                  err_msg=f"observer {t} sector {s} dir {direction}")
                    assert surf[b].view(np.uint32) == want_surf.view(np.uint32), (t, s, direction)
                    if s in compact:
                        cc, ce = compact[s]
                        k = t + direction * N
                        coff = rings.offsets(cc)
                        assert cc[k] == counts[b]
                        np.testing.assert_array_equal(ce[coff[k]:coff[k] + cc[k] - 1],
                                                      events[off[b]:off[b] + counts[b] - 1])


def test_observer_walk_at_1201_equals_the_full_pass():
    """main-1201 (235 real + 165 zero-tail steps per band): for a few
    observers the 360 band surfaces, summed in sector order, equal the
    observer's TVS from the batched pass, and three sectors' events equal
    the full-sector export's."""
    wl = demgen.main_1201()
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        e.run(0, 180, "batched")
        tvs = e.surfaces()
        N = e.n_points
        compact = {s: e.compact_rings(s) for s in (3, 100, 171)}
        for t in (0, N // 2 + 17, N - 1, 12345):
            print(f"main-1201 observer {t}", flush=True)
            counts, surf, events = e.observer_walk(t)
            assert _sum_in_sector_order(surf).view(np.uint32) == tvs[t].view(np.uint32)
            assert np.all(counts % 2 == 0)
            off = rings.offsets(counts)
            for s, (cc, ce) in compact.items():
                coff = rings.offsets(cc)
                for direction in (0, 1):
                    b, k = 2 * s + direction, t + direction * N
                    assert cc[k] == counts[b]
                    np.testing.assert_array_equal(ce[coff[k]:coff[k] + cc[k] - 1],
                                                  events[off[b]:off[b] + counts[b] - 1])
        # A sector range: the same bands.
        c2, s2, _ = e.observer_walk(12345, 40, 20)
        c_all, s_all, _ = e.observer_walk(12345)
        np.testing.assert_array_equal(c2, c_all[80:120])
        np.testing.assert_array_equal(s2.view(np.uint32), s_all[80:120].view(np.uint32))


def test_observer_walk_refusals():
    wl = demgen.small_201(seed=4)
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight) as e:
        e.upload(wl.elevations)
        with pytest.raises(_lib.TvsError) as err:
            e.observer_walk(0)                  # no band data yet
        assert err.value.status == _lib.ERR_STATE
        e.compute_bands(0, 180)
        with pytest.raises(_lib.TvsError) as err:
            e.observer_walk(e.n_points)
        assert err.value.status == _lib.ERR_INVALID_ARGUMENT
        with pytest.raises(_lib.TvsError) as err:
            e.observer_walk(0, 170, 20)
        assert err.value.status == _lib.ERR_INVALID_ARGUMENT
