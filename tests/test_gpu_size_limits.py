"""DEMs at the size limit of the second DEM copy, through the C ABI.

A DEM of up to 2^24 points also holds a copy of its heights starting at an
odd element, so the two-point walk reads every dword aligned; a larger DEM
has no copy, and the steps whose pair starts at an odd element read a dword
at 2 mod 4.  4096 x 4096 is exactly 2^24 points (copy), 4097 x 4097 is one
row and column more (no copy).  Both at 30 m with the default line of sight
(w div 3 points, 1365), three sectors (along the rows, the diagonal, along
the columns) against the oracle bit for bit: the two-point walk with the f16
form of the heights and with the int16 decode, and the one-point walk
(16-bit typed loads, which the copy does not concern)."""
import os
import sys

import numpy as np
import pytest

import tvs_amd

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
SECTORS = [0, 46, 91]
SHAPES = [(64, 8, True, False, False, True, 1),     # two points per lane, both directions
          (128, 8, True, False, False, True, 0)]    # one point per lane, both directions


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def heights(w: int) -> np.ndarray:
    """Noise of 400 m on a ramp of 2500 m across the columns: heights in
    [0, 2900], so the DEM also has its f16 form, with visibility that
    changes along every band."""
    rng = np.random.default_rng(w)
    ramp = np.linspace(0.0, 2500.0, w)[None, :]
    return (rng.integers(0, 400, size=(w, w)) + ramp).astype(np.int16).reshape(-1)


@pytest.mark.parametrize("w", [4096, 4097])
def test_second_copy_limit_bit_exact(oracle, w):
    spacing = 30.0
    los = oracle.default_max_line_of_sight(w, spacing)
    elev = heights(w)
    note(f"size limit {w}x{w}: oracle band data + walk, sectors {SECTORS}")
    dem = oracle.dem_new(w, spacing, los)
    deltas, dists = oracle.band_data_all(dem, SECTORS)
    want, _ = oracle.walk_sectors(dem, elev.astype(np.float32), 1.65, deltas, dists,
                                  n_threads=THREADS)
    with tvs_amd.Engine(w, spacing, los, 5.0, 1.65, 0) as e:
        e.upload(elev)
        assert e.half_heights() == (True, True)
        for i, s in enumerate(SECTORS):
            e.compute_bands(s, 1)
            d, x = e.get_band(s)
            np.testing.assert_array_equal(d, deltas[i], err_msg=f"{w} deltas {s}")
            np.testing.assert_array_equal(x.view(np.uint32), dists[i].view(np.uint32),
                                          err_msg=f"{w} distances {s}")
        for on in (True, False):
            e.set_half_heights(on)
            for shape in SHAPES:
                note(f"size limit {w}x{w}: half heights {on}, shape {shape}")
                e.set_walk_shape(*shape)
                e.reset_surfaces()
                for s in SECTORS:
                    e.run(s, 1, "batched")
                np.testing.assert_array_equal(e.surfaces().view(np.uint32),
                                              want.view(np.uint32),
                                              err_msg=f"{w} half {on} shape {shape}")
