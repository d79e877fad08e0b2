#!/usr/bin/env python3
"""Band-of-sight throughput benchmark (one MI355X).

One step = one full total-viewshed pass over the workload: all 180 sectors,
forward and backward bands of every computable point, per-point cumulative
surfaces.  The DEM (int16) and the 180 sectors' band data are resident in
HBM before timing starts; band data computation (device sort + compile) is
timed separately and reported as `bands_ms`.

Prints one JSON line.  `value` is point-steps/s of the batched method
(int16 heights) counted in executed steps: the batched walk skips each
band's zero-delta tail (its angle is the same at every step and its weight
is +0, so the tail changes no surface), so only the 2B - bs/2 real steps per
band are counted (`nominal_value` counts all B - 1, as the reference loop
executes them).  The plain one-sector-per-launch FP32 method executes every
step and is reported under `plain`, with the TVS relative error between the
two.  `cpu_baseline` times the oracle's literal restatement of the reference
CPU loop (every step) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "total-viewsheds_amd"))

METRIC = ("band-of-sight point-steps/s on 1201² DEM; batched vs plain at "
          "≤1e-4 TVS rel err")
HBM_PEAK_GBPS = 8000.0
# Vector-instruction issue ceiling of MI355X: 256 CUs x 4 SIMDs, each SIMD
# issuing at most two wave64 VALU instructions per four clocks (the
# SQ_ACTIVE_INST_VALU2 counter counts the quad-cycles in which a SIMD issued
# two), at the 2.4 GHz peak engine clock; in lane-operations (64 per
# wave64 instruction): 78.6e12/s.  A single-issue stream (one wave64 VALU
# instruction per four clocks) reaches half of it.
SIMDS = 256 * 4
CLOCK_HZ = 2.4e9
VALU_ISSUE_PEAK_LANE_OPS = SIMDS * (2.0 / 4.0) * CLOCK_HZ * 64
# Vector instructions every real band-step of the batched walk's unrolled
# loop needs: the height difference, the divide (4 operations, or 2 with
# the verified short divide), sub (curvature), compare -- plus one
# conversion where two points per lane share an int16 dword load and the
# DEM has no f16 form (with it, one mixed-precision FMA converts the f16
# height and subtracts the observer's; one point per lane loads with a typed
# load that converts) (tests/test_codegen.py holds the compiled loops to
# these counts).  The masked surface add and maximum move run only where a
# lane of the wave sees the point (~2 % of a lane's band-steps, ~35 % of a
# wave's at 1201 x 1201) and are not counted, so `achieved` is the work the
# walk cannot skip.
VALU_PER_STEP = {"split": 7, "short": 5}

# --no-half-heights: walk int16 DEMs with the int16 decode only.
HALF_HEIGHTS = True


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name: str):
    from tvs_amd import demgen
    return {"main-1201": demgen.main_1201, "small-201": demgen.small_201,
            "large-3601": demgen.large_3601}[name]()


def time_method(e, method: str, steps: int, warmup: int):
    """`steps` passes issued back to back between two stream waits; the walk
    kernel's duration of each (HIP events on the engine's stream) comes from
    the engine's run history afterwards (the last 64 at most)."""
    for _ in range(warmup):
        e.run(0, 180, method)
    e.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.run(0, 180, method)
    e.sync()
    wall = time.perf_counter() - t0
    kernel_ms = e.walk_history(steps)
    return wall * 1000.0 / steps, float(np.mean(kernel_ms)), e.timing()


def cpu_baseline(wl, sectors: int, threads: int, keep_surfaces: bool = False,
                 fp64: bool = False):
    """The oracle's walk on a sample of sectors.  With `keep_surfaces` also
    its cumulative f32 surfaces; with `fp64` as well the FP64 oracle walk over
    the same band data: (tvs64, visibility flips, point-steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor
    dem = O.dem_new(wl.width, wl.spacing, wl.max_line_of_sight)
    picked = list(range(0, 180, max(1, 180 // sectors)))[:sectors]
    t0 = time.perf_counter()
    if threads == 1:
        deltas, dists = O.band_data_all(dem, picked)
    else:   # one sector per task; the oracle's C code runs outside the GIL
        with ThreadPoolExecutor(threads) as ex:
            parts = list(ex.map(lambda s: O.band_data_all(dem, [s]), picked))
        deltas = np.concatenate([p[0] for p in parts])
        dists = np.concatenate([p[1] for p in parts])
    bands_s = time.perf_counter() - t0
    elev = wl.elevations.astype(np.float32)
    t0 = time.perf_counter()
    surfaces, _ = O.walk_sectors(dem, elev, wl.observer_height, deltas, dists,
                                 n_threads=threads)
    dt = time.perf_counter() - t0
    steps = len(picked) * 2 * dem.computable_points_count * dem.n_deltas
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    res = {"value": steps / dt, "unit": "point-steps/s", "cores": threads,
            "kind": "port",
            "bands_ms_per_sector": bands_s * 1000.0 / len(picked),
            "bands_s": bands_s, "walk_s": dt,
            "walk_s_180_sectors": dt * 180.0 / len(picked),   # extrapolated when sampled
            "sample": ((f"{len(picked)} of 180 sectors (every {180 // len(picked)}th)"
                        if len(picked) < 180 else "all 180 sectors") + f" of "
                       f"{wl.name}, walk only (every step, zero tail included), oracle C "
                       f"restatement of the reference CPU loop, {threads} thread(s), "
                       f"{dt:.1f} s on {cpu}; band data (axes sorts + compile) timed "
                       f"separately, {threads} thread(s)")}
    if fp64:
        t0 = time.perf_counter()
        f64 = O.walk_sectors_f64(dem, elev, wl.observer_height, deltas, dists,
                                 n_threads=threads)
        res["fp64_walk_s"] = time.perf_counter() - t0
        return res, surfaces, f64
    return (res, surfaces) if keep_surfaces else res


def rel_err_stats(got: np.ndarray, ref: np.ndarray) -> dict:
    err = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-300)
    return {"max": float(np.max(err)), "p99": float(np.percentile(err, 99)),
            "median": float(np.median(err))}


def parse_shape(text: str) -> tuple:
    """B,U,X,P,S,D,T (see --shape) as the arguments of set_walk_shape."""
    return tuple(int(v) for v in text.split(","))


def measure_large(steps: int, warmup: int, cpu: bool = False, shape_arg=None) -> dict:
    """3601 x 3601 at 30 m, LoS 40 km (1333 steps per band), int16.  With
    `cpu`, also the oracle walk on one core over every 10th sector
    (extrapolated to 180 by the executed-step rate) and on every host core
    over all 180 sectors."""
    import tvs_amd
    wl = workload("large-3601")
    log(f"large tile {wl.width}x{wl.width}: bands + tune + walk")
    with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                        wl.observer_height, 0) as e:
        g = e.geometry
        e.upload(wl.elevations)
        e.set_half_heights(HALF_HEIGHTS)
        e.compute_bands(0, 180)
        bands_ms = e.timing().bands_ms
        if shape_arg:
            e.set_walk_shape(*shape_arg)
            shape = e.walk_shape()
        else:
            shape, _ = e.autotune(0, 180, reps=3)
        ms, kernel_ms, tm = time_method(e, "batched", steps, warmup)
        e.compact_rings(45)                       # allocates; another sector
        t0 = time.perf_counter()
        counts, events = e.compact_rings(46)
        ring_ms = (time.perf_counter() - t0) * 1000.0
        assert np.all(counts % 2 == 0), "ring counts are even"
    ps = 180 * 2 * g.computable_points_count * g.real_steps
    nominal = 180 * 2 * g.computable_points_count * g.band_deltas_size
    log(f"large: {ms:.2f} ms/step, {ps / ms * 1e3:.3e} executed point-steps/s; ring export "
        f"of one sector {ring_ms:.1f} ms ({events.size} events)")
    res = {"workload": wl.name, "value": ps / (ms / 1000.0), "unit": "point-steps/s",
           "nominal_value": nominal / (ms / 1000.0),
           "ms_per_step": ms, "kernel_ms": kernel_ms, "reduce_ms": tm.reduce_ms,
           "point_steps_per_step": ps, "nominal_point_steps_per_step": nominal,
           "bands_ms": bands_ms,
           "ring_export_ms_one_sector": ring_ms, "ring_events_one_sector": int(events.size),
           "gather_GBps_l2_served": ps * 2 / (kernel_ms / 1000.0) / 1e9,
           "walk_shape": {k: shape[k] for k in ("block", "unroll", "xcd_contiguous", "directional", "pipeline",
                                                "persistent", "both_directions", "two_points")}}
    if cpu:
        log("large tile: CPU oracle walk, 1 thread, every 10th sector")
        res["cpu_one_core"] = cpu_baseline(wl, 18, 1)
        cores = min(16, len(os.sched_getaffinity(0)))
        log(f"large tile: CPU oracle walk, {cores} threads, all 180 sectors")
        res["cpu_all_cores"] = cpu_baseline(wl, 180, cores)
    return res


def cpu_tiles(tiles, sectors: int, threads: int, gpu_surfaces=None) -> dict:
    """The oracle walk over `sectors` sectors of every tile, summed; with
    `gpu_surfaces` (180 sectors only) each tile's TVS is compared with it."""
    runs = [cpu_baseline(t, sectors, threads, keep_surfaces=True) for t in tiles]
    parts = [r[0] for r in runs]
    walk_s = sum(p["walk_s"] for p in parts)
    steps = sum(p["value"] * p["walk_s"] for p in parts)
    res = {"value": steps / walk_s, "unit": "point-steps/s", "cores": threads, "kind": "port",
            "bands_s": sum(p["bands_s"] for p in parts), "walk_s": walk_s,
            "walk_s_180_sectors": sum(p["walk_s_180_sectors"] for p in parts),
            "sample": (f"{'all 180 sectors' if sectors >= 180 else f'{sectors} of 180 sectors (every {180 // sectors}th)'}"
                       f" of each of the {len(tiles)} tiles of tiles-64, walk only (every "
                       f"step, zero tail included), oracle C restatement of the reference CPU "
                       f"loop, {threads} thread(s), {walk_s:.1f} s; band data timed separately")}
    if gpu_surfaces is not None and sectors >= 180:
        res["tvs_bit_exact_vs_oracle"] = all(
            np.array_equal(g.view(np.uint32), r[1].view(np.uint32))
            for g, r in zip(gpu_surfaces, runs))
    return res


def measure_tiles(steps: int, warmup: int, cpu: bool = False, shape_arg=None) -> dict:
    """64 uneven tiles (sides from {257, 401, 601, 801, 1201}, 30/90 m), all
    180 sectors of all tiles in one walk launch + one reduction, replayed
    from a hipGraph.  With `cpu`, also the oracle walk on one core over
    every 20th sector of every tile and on every host core over all 180."""
    import tvs_amd
    from tvs_amd import demgen
    tiles = demgen.tiles_64()
    log(f"tiles-64: {sum(t.width * t.width for t in tiles)} DEM points in {len(tiles)} tiles")
    with tvs_amd.Batch([(t.width, t.spacing, t.max_line_of_sight, t.rings_per_km,
                         t.observer_height) for t in tiles]) as b:
        for i, t in enumerate(tiles):
            b.upload(i, t.elevations)
        b.set_half_heights(HALF_HEIGHTS)
        b.compute_bands()
        bands_ms = b.timing().bands_ms
        if shape_arg:
            b.set_walk_shape(*shape_arg)
            shape = dict(zip(("block", "unroll", "xcd_contiguous", "pipeline", "persistent",
                              "both_directions", "two_points"), shape_arg))
            shape["xcd_contiguous"], shape["directional"] = bool(shape_arg[2]), shape_arg[2] == 2
        else:
            shape, _ = b.autotune(reps=3)
        log(f"tiles-64 tuned shape: block {shape['block']}, unroll {shape['unroll']}, "
            f"xcd {shape['xcd_contiguous']}, directional {shape['directional']}, both directions {shape['both_directions']}, "
            f"two points {shape['two_points']}")
        for _ in range(warmup):
            b.run(True)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(steps):                 # replays back to back
            b.run(True)
        b.sync()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        graph_ms = [b.timing().walk_ms]        # the last replay (walk + reduction)
        nominal = b.timing().point_steps
        ps = sum(180 * 2 * g.computable_points_count * g.real_steps for g in b.geometry)
        b.reset_surfaces()
        b.run(False)                         # the launch pair without the graph
        tm = b.timing()
        surfaces = [b.surfaces(i) for i in range(len(tiles))]
    log(f"tiles-64: {ms:.2f} ms/step, {ps / ms * 1e3:.3e} executed point-steps/s")
    res = {"workload": "tiles-64", "value": ps / (ms / 1000.0), "unit": "point-steps/s",
            "nominal_value": nominal / (ms / 1000.0),
            "ms_per_step": ms, "graph_ms": float(np.mean(graph_ms)),
            "kernel_ms": tm.walk_ms, "reduce_ms": tm.reduce_ms,
            "point_steps_per_step": ps, "nominal_point_steps_per_step": nominal,
            "bands_ms": bands_ms,
            "tiles": len(tiles),
            "sides": sorted({t.width for t in tiles}),
            "walk_shape": {k: shape[k] for k in ("block", "unroll", "xcd_contiguous", "directional",
                                                 "pipeline", "persistent", "both_directions",
                                                 "two_points")}}
    if cpu:
        log("tiles-64: CPU oracle walk, 1 thread, every 20th sector of every tile")
        res["cpu_one_core"] = cpu_tiles(tiles, 9, 1)
        cores = min(16, len(os.sched_getaffinity(0)))
        log(f"tiles-64: CPU oracle walk, {cores} threads, all 180 sectors of every tile")
        res["cpu_all_cores"] = cpu_tiles(tiles, 180, cores, surfaces)
    return res


def measure_next_rows(wl, shape: dict) -> dict:
    """The paths beside the walk, on the same workload: `.bt` write/read and
    the CLI end to end (input file to heatmap), the band-data cache for all
    180 sectors, one sector's compact ring export, the heatmap PNG.  Each
    result is checked against its source while it is timed."""
    import subprocess
    import tempfile
    import tvs_amd
    from tvs_amd import bt
    w = wl.width
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "dem.bt")
        t0 = time.perf_counter()
        bt.write(path, wl.elevations.reshape(w, w))
        res["bt_write_ms"] = (time.perf_counter() - t0) * 1000.0
        t0 = time.perf_counter()
        back = bt.read(path)
        res["bt_read_ms"] = (time.perf_counter() - t0) * 1000.0
        assert np.array_equal(back, wl.elevations), ".bt round trip"
        res["bt_bytes"] = os.path.getsize(path)
        state = os.path.join(tmp, "state")
        with tvs_amd.Engine(w, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                            wl.observer_height, 0) as e:
            e.upload(wl.elevations)
            e.compute_bands(0, 180)
            e.set_walk_shape(shape["block"], shape["unroll"],
                             2 if shape["directional"] else int(shape["xcd_contiguous"]),
                             shape["pipeline"], shape["persistent"], shape["both_directions"],
                             shape["two_points"])
            t0 = time.perf_counter()
            for s in range(180):
                e.cache_save(state, s)
            res["cache_save_ms_180"] = (time.perf_counter() - t0) * 1000.0
            with tvs_amd.Engine(w, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                                wl.observer_height, 0) as e2:
                e2.upload(wl.elevations)
                t0 = time.perf_counter()
                for s in range(180):
                    e2.cache_load(state, s)
                res["cache_load_ms_180"] = (time.perf_counter() - t0) * 1000.0
                for s in range(0, 180, 7):
                    a, b = e.get_band(s), e2.get_band(s)
                    assert np.array_equal(a[0], b[0]) and np.array_equal(
                        a[1].view(np.uint32), b[1].view(np.uint32)), f"cache sector {s}"
            e.compact_rings(45)                       # allocates; another sector
            t0 = time.perf_counter()
            counts, events = e.compact_rings(46)
            res["ring_export_ms_one_sector"] = (time.perf_counter() - t0) * 1000.0
            assert np.all(counts % 2 == 0), "ring counts are even"
            bands = 2 * e.n_points
            res["ring_events_one_sector"] = int(events.size)
            res["ring_export_bytes"] = int(4 * (counts.size + events.size) + 8 * bands)
            # One observer's 360 bands, one wavefront per band; its sector-46
            # bands carry the same events as that sector's export.
            from tvs_amd import rings as ring_tools
            t_obs = e.n_points // 2 + 17
            e.observer_walk(t_obs)                    # allocates
            t0 = time.perf_counter()
            oc, _, oev = e.observer_walk(t_obs)
            res["observer_walk_ms_180_sectors"] = (time.perf_counter() - t0) * 1000.0
            coff, ooff = ring_tools.offsets(counts), ring_tools.offsets(oc)
            for direction in (0, 1):
                k, b = t_obs + direction * e.n_points, 2 * 46 + direction
                assert counts[k] == oc[b] and np.array_equal(
                    events[coff[k]:coff[k] + counts[k] - 1], oev[ooff[b]:ooff[b] + oc[b] - 1]), \
                    "observer walk against the sector export"
            res["ring_fixed_slot_bytes"] = int(4 * bands * tvs_amd.ring_count_per_band(
                wl.rings_per_km, wl.max_line_of_sight))
            e.run(0, 180, "batched")
            e.sync()
            # Host to host, PCIe included: upload the DEM, band data for 180
            # sectors, the walk, surfaces back (warm engine and shape); the
            # fastest of three passes, since one pass swings by ~20 % with
            # the host's load.
            h2h = []
            for _ in range(3):
                t0 = time.perf_counter()
                e.upload(wl.elevations)
                e.compute_bands(0, 180)
                e.reset_surfaces()
                e.run(0, 180, "batched")
                surf = e.surfaces()
                h2h.append((time.perf_counter() - t0) * 1000.0)
            res["host_to_host_ms"] = min(h2h)
            T = e.geometry.tvs_width
            surf = surf.reshape(T, T)
            png = os.path.join(tmp, "heatmap.png")
            t0 = time.perf_counter()
            bt.heatmap_png(png, surf)
            res["png_ms"] = (time.perf_counter() - t0) * 1000.0
            res["png_bytes"] = os.path.getsize(png)
        cli = os.path.join(ROOT, "total-viewsheds_amd", "total-viewsheds")
        out_dir = os.path.join(tmp, "cli")
        os.makedirs(out_dir)
        t0 = time.perf_counter()
        subprocess.run([cli, "--input", path, "--max-line-of-sight", str(wl.max_line_of_sight),
                        "--rings-per-km", str(wl.rings_per_km), "--observer-height",
                        str(wl.observer_height), "--scale", str(wl.spacing), "--output-dir",
                        out_dir, "--no-cache"], check=True, capture_output=True, timeout=300)
        res["cli_end_to_end_ms"] = (time.perf_counter() - t0) * 1000.0
        assert os.path.getsize(os.path.join(out_dir, "heatmap.png")) > 0
    log(f"next rows: .bt write {res['bt_write_ms']:.1f} ms / read {res['bt_read_ms']:.1f} ms, "
        f"cache save {res['cache_save_ms_180']:.1f} ms / load {res['cache_load_ms_180']:.1f} ms "
        f"(180 sectors), ring export {res['ring_export_ms_one_sector']:.1f} ms "
        f"({res['ring_events_one_sector']} events), one observer's 360 bands "
        f"{res['observer_walk_ms_180_sectors']:.2f} ms, PNG {res['png_ms']:.1f} ms, "
        f"host to host {res['host_to_host_ms']:.1f} ms, "
        f"CLI end to end {res['cli_end_to_end_ms']:.0f} ms")
    return res


def band_paths(wl_name: str) -> int:
    """Band data for all 180 sectors by both device paths -- selection
    (default) and the two full stable radix sorts per sector -- checked
    equal, with their device times."""
    import tvs_amd
    wl = workload(wl_name)
    out = {"workload": wl.name}
    bands = {}
    for method in ("select", "sort"):
        with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                            wl.observer_height, 0) as e:
            e.upload(wl.elevations)
            e.set_band_method(method)
            e.compute_bands(0, 180)
            out[f"{method}_ms"] = e.timing().bands_ms
            out[f"{method}_sorted_sectors"] = e.bands_sorted_count()
            bands[method] = [e.get_band(s) for s in range(180)]
    out["equal"] = all(np.array_equal(a[0], b[0]) and
                       np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
                       for a, b in zip(bands["select"], bands["sort"]))
    log(f"band data, 180 sectors: selection {out['select_ms']:.2f} ms, "
        f"sorts {out['sort_ms']:.1f} ms, equal {out['equal']}")
    print(json.dumps(out), flush=True)
    return 0 if out["equal"] else 1


def copy_calibration(mib: int) -> int:
    """A plain device-to-device copy of known size (the runtime's copy
    kernel), to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
    return 0


def _counter_rows(path: str):
    import csv
    with open(path) as f:
        for row in csv.DictReader(f):
            yield row


def summarize_pmc(fetch_csv: str, write_csv: str, copy_fetch_csv: str,
                  copy_write_csv: str, copy_bytes: int, out_path: str) -> int:
    """HBM bytes per launch of the batched walk from two counter runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the same
    counter reads for a copy of known size."""
    def per_kernel(path, counter):
        vals = {}
        for r in _counter_rows(path):
            if r.get("Counter_Name") != counter:
                continue
            name = r.get("Kernel_Name", "")
            vals.setdefault(name, []).append(float(r["Counter_Value"]))
        return vals

    def pick(vals, needle):
        hits = [v for k, v in vals.items() if needle in k]
        if not hits:
            raise SystemExit(f"no {needle} dispatch in counters")
        return float(np.mean(hits[0]))

    kib = copy_bytes / 1024.0
    fetch_scale = pick(per_kernel(copy_fetch_csv, "FETCH_SIZE"), "copyBuffer") / kib
    write_scale = pick(per_kernel(copy_write_csv, "WRITE_SIZE"), "copyBuffer") / kib
    fetch = pick(per_kernel(fetch_csv, "FETCH_SIZE"), "walk_tiles_kernel")
    write = pick(per_kernel(write_csv, "WRITE_SIZE"), "walk_tiles_kernel")
    res = {"kernel": "walk_tiles_kernel (batched walk, main-1201)",
           "fetch_kib_raw": fetch, "write_kib_raw": write,
           "copy_fetch_scale": fetch_scale, "copy_write_scale": write_scale,
           "hbm_bytes_per_launch": (fetch / fetch_scale + write / write_scale) * 1024.0}
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="main-1201",
                    choices=["main-1201", "small-201", "large-3601"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain method")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default walk shape")
    ap.add_argument("--shape", metavar="B,U,X,P,S,D,T",
                    help="walk shape instead of tuning: points per item, unroll, "
                         "item order (0 round-robin, 1 XCD-contiguous, 2 directional), "
                         "pipeline, persistent, both directions per lane, two points per "
                         "lane (0, 1, or 2 = two pairs from adjacent rows) "
                         "(e.g. 64,8,1,0,0,1,1)")
    ap.add_argument("--no-f32", action="store_true", help="skip the f32-heights batched run")
    ap.add_argument("--no-half-heights", action="store_true",
                    help="walk int16 DEMs without their f16 form (int16 decode only)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the large-3601 and tiles-64 figures")
    ap.add_argument("--cpu-sectors", type=int, default=36)
    ap.add_argument("--cpu-large", action="store_true",
                    help="also time the CPU oracle on the 3601x3601 tile (a few minutes)")
    ap.add_argument("--cpu-tiles", action="store_true",
                    help="also time the CPU oracle on the 64 tiles (about a minute)")
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--only", choices=["large-3601", "tiles-64"],
                    help="only time this workload (as in the bench line's `extra`) and print "
                         "its JSON")
    ap.add_argument("--band-paths", action="store_true",
                    help="only time band data by selection and by full sorts, checked equal")
    ap.add_argument("--copy-calibration", type=int, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    ap.add_argument("--summarize-pmc", nargs=6,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV",
                             "COPY_BYTES", "OUT_JSON"),
                    help="turn counter CSVs into HBM bytes per walk launch")
    args = ap.parse_args()
    global HALF_HEIGHTS
    HALF_HEIGHTS = not args.no_half_heights
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.band_paths:
        return band_paths(args.workload)
    if args.only:
        fn = measure_large if args.only == "large-3601" else measure_tiles
        print(json.dumps(fn(args.steps, args.warmup,
                            shape_arg=parse_shape(args.shape) if args.shape else None)),
              flush=True)
        return 0
    if args.summarize_pmc:
        a = args.summarize_pmc
        return summarize_pmc(a[0], a[1], a[2], a[3], int(a[4]), a[5])
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs on a single MI355X; "
              "sharding over several GPUs is not built", file=sys.stderr)
        return 2

    import tvs_amd
    wl = workload(args.workload)
    log(f"workload {wl.name}: {wl.width}x{wl.width} @ {wl.spacing} m, "
        f"los {wl.max_line_of_sight} m")
    e = tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                       wl.observer_height, 0)
    g = e.geometry
    e.upload(wl.elevations)
    e.set_half_heights(HALF_HEIGHTS)
    t0 = time.perf_counter()
    e.compute_bands(0, 180)
    bands_wall_ms = (time.perf_counter() - t0) * 1000.0
    bands_ms = e.timing().bands_ms
    log(f"band data for 180 sectors: {bands_ms:.1f} ms device, {bands_wall_ms:.1f} ms wall, "
        f"{e.bands_sorted_count()} sectors through the sorts")
    # The plain walk executes every step of every band (the reference's loop);
    # the batched walk executes the real steps and skips the zero tail.
    point_steps = 180 * 2 * g.computable_points_count * g.band_deltas_size
    real_point_steps = 180 * 2 * g.computable_points_count * g.real_steps

    if args.shape:
        e.set_walk_shape(*parse_shape(args.shape))
    shape = e.walk_shape()
    if not args.no_tune and not args.shape:
        shape, table = e.autotune(0, 180, reps=3)
        log(f"autotuned walk shape: block {shape['block']}, unroll {shape['unroll']}, "
            f"xcd-contiguous {shape['xcd_contiguous']}, directional {shape['directional']}, pipeline {shape['pipeline']}, "
            f"persistent {shape['persistent']}, both directions {shape['both_directions']}, "
            f"two points {shape['two_points']} "
            f"({shape['ms']:.2f} ms)")
        for row in sorted(table, key=lambda r: r["ms"]):
            log(f"  shape {row['block']:4d} x {row['unroll']:2d} xcd={int(row['xcd_contiguous'])} "
                f"pipe={int(row['pipeline'])} pers={int(row['persistent'])} "
                f"pair={int(row['both_directions'])} two={int(row['two_points'])}: "
                f"{row['ms']:.3f} ms")

    plain = None
    if not args.no_plain:
        log("timing plain (FP32 heights, one sector per launch)")
        p_ms, p_kernel_ms, _ = time_method(e, "plain", args.steps, args.warmup)
        plain = {"value": point_steps / (p_ms / 1000.0), "ms_per_step": p_ms,   # every step
                 "kernel_ms": p_kernel_ms, "launches_per_step": 180,
                 "achieved_GBps": point_steps * 4 / (p_kernel_ms / 1000.0) / 1e9}
        log(f"plain: {p_ms:.2f} ms/step, {plain['value']:.3e} point-steps/s")

    log("timing batched (int16 heights, all sectors in one launch)")
    b_ms, b_kernel_ms, tm = time_method(e, "batched", args.steps, args.warmup)
    _, short_sectors = e.short_divide()
    value = real_point_steps / (b_ms / 1000.0)
    log(f"batched: {b_ms:.2f} ms/step (walk kernel {b_kernel_ms:.2f} ms, reduce "
        f"{tm.reduce_ms:.2f} ms), {value:.3e} executed point-steps/s "
        f"({point_steps / (b_ms / 1000.0):.3e} nominal)")

    batched_f32 = None
    if not args.no_f32:
        # f32 heights that stay f32 on the device: every height + 0.25 m
        # (integer-valued f32 uploads are stored as int16).  Four-operation
        # divide, no short divide, one point per lane; checked against the
        # plain walk on the same heights.
        log("timing batched with f32 heights (each + 0.25 m)")
        with tvs_amd.Engine(wl.width, wl.spacing, wl.max_line_of_sight, wl.rings_per_km,
                            wl.observer_height, 0) as ef:
            ef.upload(wl.elevations.astype(np.float32) + np.float32(0.25))
            ef.compute_bands(0, 180)
            f_shape, _ = ef.autotune(0, 180, reps=3)   # f32 heights: their own best shape
            f_ms, f_kernel_ms, _ = time_method(ef, "batched", args.steps, args.warmup)
            ef.reset_surfaces()
            ef.run(0, 180, "batched")
            f_b = ef.surfaces()
            ef.reset_surfaces()
            ef.run(0, 180, "plain")
            f_equal = bool(np.array_equal(f_b.view(np.uint32), ef.surfaces().view(np.uint32)))
            f_fast = ef.fast_divide()
        batched_f32 = {"value": real_point_steps / (f_ms / 1000.0), "ms_per_step": f_ms,
                       "heights": "main-1201 + 0.25 m, f32",
                       "walk_shape": {k: f_shape[k] for k in ("block", "unroll", "both_directions",
                                                              "two_points")},
                       "nominal_value": point_steps / (f_ms / 1000.0),
                       "kernel_ms": f_kernel_ms, "fast_divide": f_fast,
                       "bit_exact_vs_plain": f_equal,
                       "achieved_GBps": real_point_steps * 4 / (f_kernel_ms / 1000.0) / 1e9}
        log(f"batched f32: {f_ms:.2f} ms/step, {batched_f32['value']:.3e} point-steps/s, "
            f"four-operation divide {f_fast}, bit-exact vs plain {f_equal}")

    # Correctness of the timed configuration: one fresh pass of each method.
    e.reset_surfaces()
    e.run(0, 180, "batched")
    tvs_b = e.surfaces()
    tvs_b64 = e.surfaces_f64()
    e.reset_surfaces()
    e.run(0, 180, "plain")
    tvs_p = e.surfaces()
    half_heights = all(e.half_heights())
    e.close()
    rel_err = float(np.max(np.abs(tvs_b.astype(np.float64) - tvs_p) / tvs_p))
    rel_err64 = float(np.max(np.abs(tvs_b64 - tvs_p) / tvs_b64))
    bit_exact = bool(np.array_equal(tvs_b.view(np.uint32), tvs_p.view(np.uint32)))

    gather_bytes = real_point_steps * 2     # one int16 height per executed step, L2-served
    per_step = VALU_PER_STEP["short"] if short_sectors == 180 else VALU_PER_STEP["split"]
    if shape["two_points"] and not half_heights:
        per_step += 1   # the two int16 heights of a dword are converted by one instruction each
    lane_ops = real_point_steps * per_step          # VALU lane-operations per launch
    valu_lane_ops = lane_ops / (b_kernel_ms / 1000.0)
    traffic = None
    prof = os.path.join(ROOT, "profiles", "walk_tiles_main-1201_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    # Issue counters of the timed shape (rocprofv3 --pmc, one profiled run of
    # this workload): every vector instruction issued, the masked updates
    # included, and how busy the texture units are.
    issue = None
    prof = os.path.join(ROOT, "profiles", "walk_tiles_main-1201_issue.json")
    timed_shape = (shape["block"], shape["unroll"], bool(shape["xcd_contiguous"]),
                   bool(shape["directional"]), bool(shape["pipeline"]),
                   bool(shape["persistent"]), bool(shape["both_directions"]),
                   int(shape["two_points"]))
    if (os.path.exists(prof) and wl.name == "main-1201"
            and timed_shape == (64, 8, True, False, False, False, True, 1)):
        with open(prof) as f:
            pj = json.load(f)
        issue = {"valu_per_band_step": pj["valu_per_band_step"],
                 "issued_frac": pj["issued_frac_of_dual_issue_ceiling"],
                 "dual_issue_quad_cycle_frac": pj["dual_issue_quad_cycle_frac"],
                 "texture_units_busy_frac": pj["texture_units_busy_frac"],
                 "source": "profiles/walk_tiles_main-1201_issue.json"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "dem": f"{wl.width}x{wl.width} int16, {wl.spacing:g} m spacing, seeded "
                   "spectral fractal" + (" + ridges" if wl.name != "small-201" else ""),
            "max_line_of_sight_m": wl.max_line_of_sight,
            "observer_height_m": wl.observer_height,
            "rings_per_km": wl.rings_per_km,
            "sectors": 180,
            "method": "batched",
            "heights": "int16",
            "point_steps_per_step": real_point_steps,
            "nominal_point_steps_per_step": point_steps,
            "steps_per_band": g.band_deltas_size,
            "real_steps_per_band": g.real_steps,
            "walk_shape": {k: shape[k] for k in ("block", "unroll", "xcd_contiguous", "directional",
                                                 "pipeline", "persistent", "both_directions",
                                                 "two_points")},
        },
        # The walk is bound by vector-instruction issue, not HBM: its height
        # gathers hit L2 (99.4 % of them), and the HBM traffic it makes is
        # ~1 % of the gathered bytes.
        "roofline": {
            "bound": "valu",
            "achieved": valu_lane_ops / 1e12,
            "peak": VALU_ISSUE_PEAK_LANE_OPS / 1e12,
            "unit": "T lane-ops/s (wave64 VALU instructions x 64)",
            "frac": valu_lane_ops / VALU_ISSUE_PEAK_LANE_OPS,
            "traffic": traffic,
            "kernel": "walk_tiles_kernel<int16_t> (batched walk)",
            "kernel_ms": b_kernel_ms,
            "valu_per_band_step": per_step,
            "half_heights": half_heights,
            "short_divide_sectors": short_sectors,
            "lane_ops_per_launch": lane_ops,
            "peak_basis": "1024 SIMDs x 2 wave64 VALU instructions per 4 clocks (dual issue) "
                          "x 2.4 GHz x 64 lanes",
            "hbm": {"bytes_per_launch": traffic,
                    "achieved_GBps": (traffic / (b_kernel_ms / 1000.0) / 1e9
                                      if traffic else None),
                    "peak_GBps": HBM_PEAK_GBPS,
                    "frac": (traffic / (b_kernel_ms / 1000.0) / 1e9 / HBM_PEAK_GBPS
                             if traffic else None)},
            "l2_served_gather_bytes_per_launch": gather_bytes,
            "profiled_issue": issue,
        },
        "nominal_value": point_steps / (b_ms / 1000.0),
        "plain": plain,
        "batched_f32": batched_f32,
        "tvs_rel_err_batched_vs_plain": rel_err,
        "tvs_bit_exact_batched_vs_plain": bit_exact,
        "tvs_rel_err_f64_sum_vs_plain": rel_err64,
        "reduce_ms": tm.reduce_ms,
        "bands_ms": bands_ms,
    }
    if not args.no_extra:
        out["next_rows"] = measure_next_rows(wl, shape)
        out["extra"] = [measure_large(max(2, args.steps // 2), 1, args.cpu_large),
                        measure_tiles(max(2, args.steps // 2), 1, args.cpu_tiles)]
    if not args.no_cpu_baseline:
        log(f"CPU baseline: oracle walk over {args.cpu_sectors} sectors, "
            f"{args.cpu_threads} thread(s)")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_sectors, args.cpu_threads)
        log(f"CPU: {out['cpu_baseline']['value']:.3e} point-steps/s")
        # All 180 sectors over every host core this process may use (capped
        # at 16, the host share of one GPU), reported beside the contract's
        # single-thread sample.
        cores = min(16, len(os.sched_getaffinity(0)))
        if cores > 1:
            out["cpu_all_cores"], cpu_surf, (tvs64, flips, steps64) = cpu_baseline(
                wl, 180, cores, keep_surfaces=True, fp64=True)
            # The oracle's 180-sector TVS against the batched pass checked above.
            out["tvs_bit_exact_vs_oracle"] = bool(
                np.array_equal(cpu_surf.view(np.uint32), tvs_b.view(np.uint32)))
            # The precision form against the FP64 oracle (the same band data,
            # angles, maxima and sums in f64): the f32 walk's visibility
            # flips, and the relative error of the batched pass's f32 TVS and
            # of its f64 per-point sums.
            out["fp64_oracle"] = {
                "point_steps": steps64, "visibility_flips": flips,
                "flip_rate": flips / steps64 if steps64 else None,
                "tvs_f32_rel_err": rel_err_stats(tvs_b, tvs64),
                "tvs_f64_sum_rel_err": rel_err_stats(tvs_b64, tvs64),
                "cpu_walk_s": out["cpu_all_cores"].pop("fp64_walk_s")}
            log(f"CPU, {cores} threads: {out['cpu_all_cores']['value']:.3e} point-steps/s; "
                f"batched TVS bit-exact vs the oracle: {out['tvs_bit_exact_vs_oracle']}; "
                f"FP64 oracle: {flips} flips, f32 TVS rel err p99 "
                f"{out['fp64_oracle']['tvs_f32_rel_err']['p99']:.2e}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
