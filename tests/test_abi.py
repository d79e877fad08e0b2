"""The C ABI library: it loads, exports every entry point include/tvs.h
declares, and its host-only geometry agrees with the oracle.  No kernel is
launched here (CPU only)."""
import ctypes
import re

import numpy as np
import pytest

import tvs_amd
from tvs_amd import _lib


def declared_functions():
    text = open(tvs_amd.HEADER_PATH).read()
    return sorted(set(re.findall(r"^[a-z_ *]+?\b(tvs_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("tvs_create", "tvs_destroy", "tvs_upload_dem_f32", "tvs_upload_dem_i16",
                 "tvs_compute_bands", "tvs_set_band", "tvs_get_band", "tvs_compute_axes",
                 "tvs_run_sectors", "tvs_run_sector_with_band", "tvs_read_surfaces",
                 "tvs_read_surfaces_f64", "tvs_read_rings", "tvs_ring_overflow_count",
                 "tvs_last_error", "tvs_dem_geometry"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tvs_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_entry_points_open_marker_ranges():
    """The library calls roctx (rocprofv3 --marker-trace shows a named range
    per entry point call) and loads its runtime library."""
    import shutil
    import subprocess
    if not shutil.which("nm"):
        pytest.skip("no nm")
    syms = subprocess.run(["nm", "-D", "--undefined-only", tvs_amd.LIB_PATH],
                          capture_output=True, text=True, check=True).stdout
    assert "roctxRangePushA" in syms and "roctxRangePop" in syms
    ctypes.CDLL(tvs_amd.LIB_PATH)


def test_page_locked_memory_where_no_device_can_pin_it():
    """tvs_host_alloc(0) is NULL and tvs_host_free(NULL) succeeds anywhere.
    Without a device (this CPU-only run) every allocation is NULL and
    pinned_empty raises MemoryError, which the ring export answers with
    ordinary arrays."""
    lib = _lib.load()
    assert not lib.tvs_host_alloc(0)
    assert lib.tvs_host_free(None) == 0
    p = lib.tvs_host_alloc(64)
    if p:                                   # a device is present: nothing to check here
        assert lib.tvs_host_free(p) == 0
        pytest.skip("a device can pin memory here")
    with pytest.raises(MemoryError):
        _lib.pinned_empty(16, np.uint32)


@pytest.mark.parametrize("w,scale,los", [
    (9, 1.0, 3), (201, 100.0, 6700), (1201, 90.0, 36000), (3601, 30.0, 40000),
    (257, 30.0, 2550), (401, 90.0, 11970), (601, 30.0, 6000), (801, 90.0, 24030),
    (100, 7.5, 300), (333, 12.25, 1337), (50, 1.0, 0), (64, 3.0, 96),
])
def test_host_geometry_matches_oracle(oracle, w, scale, los):
    g = tvs_amd.dem_geometry(w, scale, los)
    d = oracle.dem_new(w, scale, los)
    assert (g.dem_width, g.dem_size, g.max_line_of_sight, g.max_los_as_points,
            g.computable_points_count, g.tvs_width, g.band_size) == \
        (d.width, d.size, d.max_line_of_sight, d.max_los_as_points,
         d.computable_points_count, d.tvs_width, d.band_size)
    assert g.band_deltas_size == d.band_size - 1
    assert g.band_samples == int(2 * d.band_size * np.sqrt(2.0))


def test_observer_origin_quirk_at_3601():
    """dem.rs:106-114 vs kernel.rs:125-126: at 3601 x 30 m with 40 km the
    computable columns start at 1334 but observers start at L = 1333."""
    g = tvs_amd.dem_geometry(3601, 30.0, 40000)
    first = next(x for x in range(3601) if np.float32(x) * np.float32(30.0) >= np.float32(40000))
    assert (g.max_los_as_points, first) == (1333, 1334)


@pytest.mark.parametrize("w,scale", [(201, 100.0), (1201, 90.0), (257, 30.0), (3601, 30.0),
                                     (1000, 0.3), (17, 1e6)])
def test_default_line_of_sight(oracle, w, scale):
    assert tvs_amd.default_max_line_of_sight(w, scale) == \
        oracle.default_max_line_of_sight(w, scale)


@pytest.mark.parametrize("rpk,los", [(5.0, 36000), (5.0, 6700), (5000.0, 3), (0.1, 2550),
                                     (7.3, 40000), (5.0, 0)])
def test_ring_count_per_band(oracle, rpk, los):
    assert tvs_amd.ring_count_per_band(rpk, los) == oracle.ring_count_per_band(rpk, los)


def test_los_too_long_status():
    with pytest.raises(tvs_amd.TvsError) as e:
        tvs_amd.dem_geometry(9, 1.0, 5)
    assert e.value.status == _lib.ERR_LOS_TOO_LONG


def test_invalid_arguments_status():
    for args in ((0, 1.0, 0), (9, 0.0, 1), (9, -1.0, 1), (50000, 1.0, 1)):
        with pytest.raises(tvs_amd.TvsError) as e:
            tvs_amd.dem_geometry(*args)
        assert e.value.status == _lib.ERR_INVALID_ARGUMENT


def test_status_strings():
    lib = _lib.load()
    for st in range(0, 11):
        assert lib.tvs_status_string(st)
    assert lib.tvs_last_error(None) is not None


def test_demgen_is_deterministic():
    from tvs_amd import demgen
    a = demgen.fractal_dem(129, 30.0, 2.4, 4000.0, seed=3, ridges=2)
    b = demgen.fractal_dem(129, 30.0, 2.4, 4000.0, seed=3, ridges=2)
    c = demgen.fractal_dem(129, 30.0, 2.4, 4000.0, seed=4, ridges=2)
    assert a.dtype == np.int16 and a.shape == (129, 129)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() == 0 and a.max() == 4000


def test_tiles_workload_shape():
    from tvs_amd import demgen
    tiles = demgen.tiles_64(count=6)
    for t in tiles:
        assert t.width in demgen.TILE_SIDES and t.spacing in (30.0, 90.0)
        assert t.max_line_of_sight == (t.width // 3) * int(t.spacing)
        assert t.elevations.size == t.width * t.width
