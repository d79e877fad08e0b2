"""Many tiles in one launch (tvs_batch), hipGraph replay, and the launch
shape tuner: every tile's surfaces equal its own single-context run and the
oracle, bit for bit."""
import os

import numpy as np
import pytest

import tvs_amd
from tvs_amd import demgen, _lib

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def small_tiles():
    """Six uneven tiles (two share width and band size with others)."""
    specs = [(129, 30.0, 1), (87, 90.0, 2), (129, 30.0, 3), (101, 30.0, 4), (87, 90.0, 5),
             (61, 90.0, 6)]
    out = []
    for w, sp, seed in specs:
        los = int((w // 3) * sp)
        out.append(demgen.Workload(f"t{seed}", w, sp, los,
                                   demgen.fractal_dem(w, sp, 2.4, 4000.0, seed).reshape(-1)))
    return out


@pytest.fixture(scope="module")
def tiles_and_truth(oracle):
    tiles = small_tiles()
    truth = []
    for t in tiles:
        dem = oracle.dem_new(t.width, t.spacing, t.max_line_of_sight)
        cum, _, _ = oracle.compute_run(dem, t.elevations.astype(np.float32), t.observer_height,
                                       t.rings_per_km, rings=False, n_threads=THREADS)
        truth.append(cum)
    return tiles, truth


def test_batch_mixed_height_types(oracle, tiles_and_truth):
    """Tiles uploaded as int16, as integer-valued f32 (stored as int16) and
    as non-integer f32 share one launch: it walks every tile's f32 heights,
    each tile's surfaces equal the oracle's."""
    tiles, truth = tiles_and_truth
    rng = np.random.default_rng(9)
    frac = {}
    for i in (1, 4):
        t = tiles[i]
        frac[i] = (t.elevations.astype(np.float64) + 0.25 + 0.5 * rng.random(t.elevations.size)
                   ).astype(np.float32)
        dem = oracle.dem_new(t.width, t.spacing, t.max_line_of_sight)
        cum, _, _ = oracle.compute_run(dem, frac[i], t.observer_height, t.rings_per_km,
                                       rings=False, n_threads=THREADS)
        frac[i] = (frac[i], cum)
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            if i in frac:
                b.upload(i, frac[i][0])
            elif i % 2:
                b.upload(i, t.elevations.astype(np.float32))
            else:
                b.upload(i, t.elevations)
        b.compute_bands()
        b.run(False)
        for i in range(len(tiles)):
            want = frac[i][1] if i in frac else truth[i]
            np.testing.assert_array_equal(b.surfaces(i).view(np.uint32), want.view(np.uint32),
                                          err_msg=f"tile {i}")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_oracle(tiles_and_truth, use_graph):
    tiles, truth = tiles_and_truth
    print(f"batch of {len(tiles)} tiles, graph={use_graph}", flush=True)
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.compute_bands()
        b.run(use_graph)
        for i in range(len(tiles)):
            np.testing.assert_array_equal(b.surfaces(i), truth[i], err_msg=f"tile {i}")
        # Replays: eager reset between them, every replay equal to the first.
        for rep in range(3):
            b.reset_surfaces()
            b.run(use_graph)
            b.sync()
            for i in range(len(tiles)):
                np.testing.assert_array_equal(b.surfaces(i), truth[i],
                                              err_msg=f"replay {rep} tile {i}")


def test_batch_shapes_agree(tiles_and_truth):
    tiles, truth = tiles_and_truth
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.compute_bands()
        for block, unroll, xcd, pipe, pair, two in [(64, 8, False, False, False, False),
                                                    (128, 16, True, False, False, False),
                                                    (512, 8, False, True, False, False),
                                                    (256, 4, True, True, False, False),
                                                    (64, 8, False, False, True, False),
                                                    (128, 4, True, False, True, False),
                                                    (128, 8, False, False, False, True),
                                                    (64, 4, True, False, True, True),
                                                    (64, 8, 2, False, True, True),
                                                    (256, 8, 2, False, False, False)]:
            print(f"batch shape {block}x{unroll} xcd={xcd} pipe={pipe} pair={pair} two={two}",
                  flush=True)
            b.set_walk_shape(block, unroll, xcd, pipe, both_directions=pair, two_points=two)
            b.reset_surfaces()
            b.run(True)
            for i in range(len(tiles)):
                np.testing.assert_array_equal(b.surfaces(i), truth[i])


def test_batch_half_heights_toggle(tiles_and_truth):
    """The small tiles' heights (0 to 4000 m) all have the f16 form; the
    batch's two-point walk with it on, off and on again (the graph is
    rebuilt each time) equals the oracle bit for bit."""
    tiles, truth = tiles_and_truth
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.compute_bands()
        b.set_walk_shape(64, 8, True, False, both_directions=True, two_points=1)
        for on in (True, False, True):
            print(f"batch half heights {on}", flush=True)
            b.set_half_heights(on)
            b.reset_surfaces()
            b.run(True)
            for i in range(len(tiles)):
                np.testing.assert_array_equal(b.surfaces(i).view(np.uint32),
                                              truth[i].view(np.uint32), err_msg=f"{on} tile {i}")


def test_autotune_keeps_results_exact(tiles_and_truth):
    tiles, truth = tiles_and_truth
    t = tiles[0]
    with tvs_amd.Engine(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                        t.observer_height) as e:
        e.upload(t.elevations)
        e.compute_bands(0, 180)
        e.run(0, 180, "batched")
        before = e.surfaces()
        best, table = e.autotune(0, 30, reps=1)
        print("autotune", best, flush=True)
        # every compiled shape in both item orders, the two-point shapes and
        # the one-point shapes walking both directions also in the
        # directional one, persistent shapes once
        assert len(table) == _lib.load().tvs_walk_shape_count() == 2 * 23 + 5 + 12
        assert best["ms"] > 0
        np.testing.assert_array_equal(e.surfaces(), before)   # surfaces preserved
        assert e.walk_shape()["block"] == best["block"]
        e.reset_surfaces()
        e.run(0, 180, "batched")
        np.testing.assert_array_equal(e.surfaces(), truth[0])


def test_batch_autotune_keeps_results_exact(tiles_and_truth):
    """tvs_batch_autotune times only the walk: the cumulative surfaces are
    unchanged, and the picked shape's run is bit-exact."""
    tiles, truth = tiles_and_truth
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.compute_bands()
        b.run(True)
        before = [b.surfaces(i) for i in range(len(tiles))]
        best, table = b.autotune(reps=1)
        print("batch autotune", best, flush=True)
        assert len(table) == _lib.load().tvs_walk_shape_count() == 2 * 23 + 5 + 12
        assert best["ms"] > 0
        for i in range(len(tiles)):
            np.testing.assert_array_equal(b.surfaces(i), before[i])
        b.reset_surfaces()
        b.run(True)
        for i in range(len(tiles)):
            np.testing.assert_array_equal(b.surfaces(i), truth[i], err_msg=f"tile {i}")


@pytest.mark.parametrize("tiny_tiles", [(0, 1, 2, 3, 4, 5), (1, 4)])
def test_batch_ieee_divide_tiles(oracle, tiles_and_truth, tiny_tiles):
    """Tiles whose heights fail the fast-divide range check (one non-zero
    height below 2^-17) walk with the IEEE divide: all of them (the launch
    takes the IEEE-only kernels) or some (those tiles take the one-step IEEE
    loop of the fast kernels).  Every tile equals the oracle."""
    tiles, truth = tiles_and_truth
    want = list(truth)
    heights = [t.elevations for t in tiles]
    for i in tiny_tiles:
        t = tiles[i]
        h = t.elevations.astype(np.float32) + np.float32(0.5)
        h[t.width * (t.width // 2) + 3] = np.float32(1e-30)
        heights[i] = h
        dem = oracle.dem_new(t.width, t.spacing, t.max_line_of_sight)
        want[i], _, _ = oracle.compute_run(dem, h, t.observer_height, t.rings_per_km,
                                           rings=False, n_threads=THREADS)
    print(f"batch with IEEE-divide tiles {tiny_tiles}", flush=True)
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, h in enumerate(heights):
            b.upload(i, h)
        b.compute_bands()
        b.run(True)
        b.sync()
        for i in range(len(tiles)):
            np.testing.assert_array_equal(b.surfaces(i).view(np.uint32),
                                          want[i].view(np.uint32), err_msg=f"tile {i}")


def test_batch_mutators_wait_for_the_running_batch(oracle, tiles_and_truth):
    """tvs_batch_run is asynchronous: an upload (or any other change) issued
    right after it, with no wait in between, must not reach the heights the
    running walk reads.  Run the graph, upload new heights to every tile at
    once, run again: the first results are the old heights' (read after the
    second run is queued), the second the new heights'."""
    tiles, truth = tiles_and_truth
    rng = np.random.default_rng(21)
    new = []
    want_new = []
    for t in tiles:
        h = rng.permutation(t.elevations)
        new.append(h)
        dem = oracle.dem_new(t.width, t.spacing, t.max_line_of_sight)
        cum, _, _ = oracle.compute_run(dem, h.astype(np.float32), t.observer_height,
                                       t.rings_per_km, rings=False, n_threads=THREADS)
        want_new.append(cum)
    print("batch: upload right after an asynchronous run", flush=True)
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.compute_bands()
        b.set_walk_shape(256, 8)          # one band per lane: the slowest shape
        b.run(True)                       # no sync before the uploads
        for i, h in enumerate(new):
            b.upload(i, h)
        first = [b.surfaces(i) for i in range(len(tiles))]
        b.reset_surfaces()
        b.run(True)
        b.set_walk_shape(64, 8, both_directions=True, two_points=True)   # no sync either
        second = [b.surfaces(i) for i in range(len(tiles))]
        for i in range(len(tiles)):
            np.testing.assert_array_equal(first[i], truth[i], err_msg=f"old heights, tile {i}")
            np.testing.assert_array_equal(second[i], want_new[i], err_msg=f"new heights, tile {i}")


def test_autotune_leaves_ring_capture_alone(tiles_and_truth):
    """With ring capture on, the tuner's timing runs record no rings (every
    shape is timed as itself, not as the capture kernel) and the rings and
    overflow count of the caller's last run stay readable and unchanged."""
    tiles, _ = tiles_and_truth
    t = tiles[0]
    with tvs_amd.Engine(t.width, t.spacing, t.max_line_of_sight, 0.5,
                        t.observer_height, 1) as e:
        e.upload(t.elevations)
        e.compute_bands(0, 180)
        e.run(17, 1, "batched")
        rings_before = e.rings(0)
        overflow_before = e.ring_overflow()
        assert overflow_before > 0          # R is small: some bands overflow
        best, table = e.autotune(0, 12, reps=1)
        print("autotune with capture on", best, flush=True)
        np.testing.assert_array_equal(e.rings(0), rings_before)
        assert e.ring_overflow() == overflow_before
        assert len(table) == _lib.load().tvs_walk_shape_count() and best["ms"] > 0


@pytest.mark.parametrize("sides,directional", [((257, 401), False), ((2897, 257), True)])
def test_batch_default_item_order_by_tile_size(sides, directional):
    """An untuned batch walks its items round-robin over the XCDs, or in the
    directional order when a tile's int16 heights take more than 16 MiB
    (2897^2 points).  Either way every tile's TVS equals its own context's
    run in the XCD-contiguous order, bit for bit; a set shape overrides the
    default."""
    elev = [np.random.default_rng(w).integers(0, 2000, w * w).astype(np.int16) for w in sides]
    geo = [(w, 30.0, (w // 3) * 30, 5.0, 1.65) for w in sides]
    want = []
    for g, h in zip(geo, elev):
        with tvs_amd.Engine(*g, 0) as e:
            e.upload(h)
            e.compute_bands(0, 180)
            e.set_walk_shape(64, 8, 1, False, False, True, 1)
            e.run(0, 180, "batched")
            want.append(e.surfaces())
    with tvs_amd.Batch(geo) as b:
        for i, h in enumerate(elev):
            b.upload(i, h)
        shape = b.walk_shape()
        print(f"batch {sides}: default shape {shape}", flush=True)
        assert shape["directional"] == directional
        assert shape["xcd_contiguous"] == directional
        b.compute_bands()
        b.run(True)
        b.sync()
        for i in range(len(sides)):
            np.testing.assert_array_equal(b.surfaces(i).view(np.uint32),
                                          want[i].view(np.uint32), err_msg=f"tile {i}")
        b.set_walk_shape(64, 8, 1, both_directions=True, two_points=1)
        s2 = b.walk_shape()
        assert s2["xcd_contiguous"] and not s2["directional"]
