"""The `total-viewsheds` CLI end to end (config.rs flags, main.rs flow) and
the band-data cache (cache.rs layout), on the 201 x 201 workload."""
import os
import subprocess

import numpy as np
import pytest

import tvs_amd
from tvs_amd import bt, demgen, _lib

from test_io import read_png_grey

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(tvs_amd.LIB_PATH), "total-viewsheds")


def png_of(surfaces, T):
    data = surfaces.reshape(T, T).astype(np.float32)
    mn, mx = data.min(), data.max()
    rng = np.float32(1.0) if mx - mn == 0 else np.float32(mx - mn)
    return np.clip((data - mn) / rng * np.float32(255.0), 0, 255).astype(np.uint8)


@pytest.fixture(scope="module")
def small(oracle):
    wl = demgen.small_201(seed=3)
    dem = oracle.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    cum, _, _ = oracle.compute_run(dem, wl.elevations.astype(np.float32), wl.observer_height,
                                   wl.rings_per_km, rings=False, n_threads=8)
    return wl, dem, cum


@pytest.mark.parametrize("method", ["hip", "hip-plain"])
def test_cli_heatmap(tmp_path, small, method):
    wl, dem, cum = small
    path = tmp_path / "dem.bt"
    bt.write(str(path), wl.elevations.reshape(wl.width, wl.width))
    state = tmp_path / "state"
    out = tmp_path / "out"
    args = [CLI, "--input", str(path), "--scale", "100", "--max-line-of-sight", "6700",
            "--compute", method, "--output-dir", str(out), "--state-dir", str(state)]
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(read_png_grey(str(out / "heatmap.png")),
                                  png_of(cum, dem.tvs_width))
    cache = state / "total-viewsheds" / "201"
    assert len(os.listdir(cache / "band_deltas")) == 180
    assert os.path.getsize(cache / "distances" / "17.bin") == 4 * dem.n_deltas
    # Second run reads the cache and draws the same picture.
    r2 = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    assert "Computing band data" not in r2.stderr
    np.testing.assert_array_equal(read_png_grey(str(out / "heatmap.png")),
                                  png_of(cum, dem.tvs_width))


def test_cli_tune_caches_the_walk_shape(tmp_path, small):
    """--tune times the walk's launch shapes on the DEM and keeps the
    fastest; the choice is written under the state directory per width,
    line of sight in points (6700 m at 100 m: 67) and height type, and the next run loads it instead of
    tuning again.  Both runs draw the oracle's picture."""
    wl, dem, cum = small
    path = tmp_path / "dem.bt"
    bt.write(str(path), wl.elevations.reshape(wl.width, wl.width))
    state = tmp_path / "state"
    out = tmp_path / "out"
    args = [CLI, "--input", str(path), "--scale", "100", "--max-line-of-sight", "6700",
            "--output-dir", str(out), "--state-dir", str(state), "--tune"]
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Tuning the walk's launch shape" in r.stderr
    shape_file = state / "total-viewsheds" / "201" / "walk_shape" / "67pts_i16.txt"
    fields = shape_file.read_text().split()
    assert len(fields) == 7 and int(fields[0]) in (64, 128, 256, 512)
    np.testing.assert_array_equal(read_png_grey(str(out / "heatmap.png")),
                                  png_of(cum, dem.tvs_width))
    r2 = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    assert "Tuning" not in r2.stderr and "Walk shape from" in r2.stderr
    np.testing.assert_array_equal(read_png_grey(str(out / "heatmap.png")),
                                  png_of(cum, dem.tvs_width))
    # A stored shape this build does not have is ignored: tuned again.
    shape_file.write_text("48 3 0 0 0 0 0\n")
    r3 = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r3.returncode == 0, r3.stderr
    assert "Tuning the walk's launch shape" in r3.stderr
    assert shape_file.read_text().split()[0] in ("64", "128", "256", "512")


def test_cli_default_line_of_sight_and_errors(tmp_path, small):
    wl, dem, _ = small
    path = tmp_path / "dem.bt"
    bt.write(str(path), wl.elevations.reshape(wl.width, wl.width))
    # main.rs:36-38: default LoS = (w div 3) * scale = 6700 m at 100 m.
    r = subprocess.run([CLI, "--input", str(path), "--scale", "100", "--no-cache",
                        "--output-dir", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "max line of sight 6700 m" in r.stderr
    r = subprocess.run([CLI, "--input", str(path), "--scale", "100",
                        "--max-line-of-sight", "20000", "--no-cache"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "longer than the maximum distance" in r.stderr
    r = subprocess.run([CLI, "--input", str(path), "--compute", "vulkan"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def test_band_cache_round_trip(tmp_path, small):
    wl, dem, cum = small
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        e.compute_bands(0, 180)
        for s in range(180):
            e.cache_save(str(tmp_path), s)
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height) as e:
        e.upload(wl.elevations)
        for s in range(180):
            e.cache_load(str(tmp_path), s)
        e.run(0, 180, "batched")
        np.testing.assert_array_equal(e.surfaces(), cum)
    # A file of another band length (another line of sight) is refused.
    f = tmp_path / "total-viewsheds" / "201" / "distances" / "5.bin"
    f.write_bytes(f.read_bytes()[:-4])
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight) as e:
        with pytest.raises(tvs_amd.TvsError) as err:
            e.cache_load(str(tmp_path), 5)
        assert err.value.status == _lib.ERR_SIZE


@pytest.fixture(scope="module")
def running_total(oracle, small):
    """compute.rs:206-214 on the oracle: the f32 running total of the
    cumulative surfaces after every sector (what the per-sector heatmap
    shows)."""
    wl, dem, _ = small
    elev = wl.elevations.astype(np.float32)
    cum = np.zeros(dem.computable_points_count, np.float32)
    total = np.zeros_like(cum)
    for s in range(180):
        oracle.compute_run(dem, elev, wl.observer_height, wl.rings_per_km, first_sector=s,
                           sector_count=1, rings=False, n_threads=8, cumulative=cum)
        total += cum
    return total


def test_compute_mirror_cache_and_heatmap(tmp_path, oracle, small, running_total):
    """The Python mirror of `Compute` (compute.rs:101-178): with a state
    directory the band data of every sector is computed once and saved in
    the reference's cache layout, then loaded by the next run; with an output
    directory heatmap.png shows the running total (compute.rs:206-214)."""
    from tvs_amd.compute import DEM, Compute
    wl, dem, cum = small
    d = DEM(wl.width, wl.spacing, wl.max_line_of_sight)
    d.elevations = wl.elevations
    state, out = tmp_path / "state", tmp_path / "out"
    c = Compute("hip", state, out, d, wl.rings_per_km, wl.observer_height, capture_rings=False)
    surfaces, _ = c.run()
    assert c.sectors_loaded == 0
    np.testing.assert_array_equal(surfaces, cum)
    np.testing.assert_array_equal(c.total_surfaces, running_total)
    np.testing.assert_array_equal(read_png_grey(str(out / "heatmap.png")),
                                  png_of(running_total, dem.tvs_width))
    cache = state / "total-viewsheds" / "201"
    deltas, dists = oracle.band_data(dem, 37)
    assert (cache / "band_deltas" / "37.bin").read_bytes() == deltas.tobytes()
    assert (cache / "distances" / "37.bin").read_bytes() == dists.tobytes()
    # The next run loads all 180 sectors; a file of another band length is
    # recomputed and rewritten.
    f = cache / "distances" / "5.bin"
    f.write_bytes(f.read_bytes()[:-4])
    c2 = Compute("hip-plain", state, None, d, wl.rings_per_km, wl.observer_height,
                 capture_rings=False)
    surfaces2, _ = c2.run()
    assert c2.sectors_loaded == 179
    np.testing.assert_array_equal(surfaces2, cum)
    assert f.read_bytes() == oracle.band_data(dem, 5)[1].tobytes()


def test_cli_heatmap_each_sector(tmp_path, small, running_total):
    """--heatmap-each-sector redraws heatmap.png after every sector from the
    running total, as compute.rs:106-112 does."""
    wl, dem, _ = small
    path = tmp_path / "dem.bt"
    bt.write(str(path), wl.elevations.reshape(wl.width, wl.width))
    out = tmp_path / "out"
    r = subprocess.run([CLI, "--input", str(path), "--scale", "100", "--max-line-of-sight",
                        "6700", "--output-dir", str(out), "--no-cache",
                        "--heatmap-each-sector"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(read_png_grey(str(out / "heatmap.png")),
                                  png_of(running_total, dem.tvs_width))
