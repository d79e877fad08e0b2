"""tvs_run_sectors_noalloc (band surfaces in the library's static device
scratch, no allocation) against tvs_run_sectors (band surfaces in a buffer
grown on demand), bit for bit, through the C ABI.

Two calls in a row share the scratch: first the 1201 x 1201 workload over
all 180 sectors (four chunks of at most 52 sectors), then a smaller DEM
(201 x 201, a sector range, a directional two-point shape) on which the
first call's band surfaces are still in the scratch; the second engine runs
twice to check that the cumulative surfaces accumulate as in
tvs_run_sectors.  Then the two refusals: ring capture, and a DEM whose one
sector of band surfaces exceeds the scratch."""
import numpy as np
import pytest

import tvs_amd
from tvs_amd import _lib, demgen

pytestmark = pytest.mark.gpu


def _engines(wl, shape=None):
    pair = []
    for _ in range(2):
        e = tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                           wl.observer_height)
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        if shape:
            e.set_walk_shape(*shape)
        pair.append(e)
    return pair


def _same(got, want, what):
    np.testing.assert_array_equal(got.view(got.dtype.str.replace("f", "u")),
                                  want.view(want.dtype.str.replace("f", "u")), err_msg=what)


def test_noalloc_equals_the_allocating_run_on_two_calls():
    wl = demgen.main_1201()
    print("noalloc: main-1201, 180 sectors", flush=True)
    ref, e = _engines(wl)
    try:
        ref.run(0, 180, "batched")
        e.run_noalloc(0, 180)
        assert e.timing().walk_launches == 4          # 52 + 52 + 52 + 24 sectors
        _same(e.surfaces(), ref.surfaces(), "main-1201 f32")
        _same(e.surfaces_f64(), ref.surfaces_f64(), "main-1201 f64")
    finally:
        ref.close()
        e.close()

    small = demgen.small_201(seed=3)
    print("noalloc: small-201, sectors 30..126, directional two-point shape", flush=True)
    ref, e = _engines(small, shape=(64, 8, 2, False, False, True, True))
    try:
        for rep in range(2):
            ref.run(30, 97, "batched")
            e.run_noalloc(30, 97)
            assert e.timing().walk_launches == 1
            _same(e.surfaces(), ref.surfaces(), f"small-201 f32, run {rep}")
            _same(e.surfaces_f64(), ref.surfaces_f64(), f"small-201 f64, run {rep}")
    finally:
        ref.close()
        e.close()


def test_noalloc_f32_heights():
    """Non-integer f32 heights (the f32 kernels, the four-operation divide):
    the same bit-exact agreement."""
    small = demgen.small_201(seed=5)
    small.elevations = small.elevations.astype(np.float32) + np.float32(0.25)
    print("noalloc: small-201 + 0.25 m, f32 heights", flush=True)
    ref, e = _engines(small)
    try:
        assert e.fast_divide()
        ref.run(0, 180, "batched")
        e.run_noalloc(0, 180)
        _same(e.surfaces(), ref.surfaces(), "f32 heights, f32 sums")
        _same(e.surfaces_f64(), ref.surfaces_f64(), "f32 heights, f64 sums")
    finally:
        ref.close()
        e.close()


def test_noalloc_refusals():
    small = demgen.small_201(seed=3)
    with tvs_amd.Engine(small.width, small.spacing, small.max_line_of_sight,
                        small.rings_per_km, small.observer_height, 1) as e:
        e.upload(small.elevations)
        e.compute_bands(0, 180)
        with pytest.raises(_lib.TvsError) as err:
            e.run_noalloc(0, 180)
        assert err.value.status == 1 and "ring capture" in str(err.value)
    # 8701 x 8701 at 1 m, LoS 2900 m: N = 2901^2, 2N x 4 bytes > 64 MiB.
    with tvs_amd.Engine(8701, 1.0, 2900) as big:
        assert big.n_points == 2901 * 2901
        with pytest.raises(_lib.TvsError) as err:
            big.run_noalloc(0, 180)
        assert err.value.status == 5 and "static scratch" in str(err.value)
