"""ctypes binding of libtvs_hip.so (the C ABI declared in include/tvs.h).

The library must have been built (``make -C total-viewsheds_amd`` or
``__graft_entry__.build()``); loading fails loudly otherwise -- there is no
CPU fallback in the engine.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libtvs_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "tvs.h")

# tvs_status
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_LOS_TOO_LONG = 2
ERR_BAND_TOO_SHORT = 3
ERR_BAND_OUT_OF_DEM = 4
ERR_SIZE = 5
ERR_STATE = 6
ERR_NO_DEVICE = 7
ERR_OUT_OF_MEMORY = 8
ERR_HIP = 9
ERR_IO = 10

# tvs_method
METHOD_PLAIN = 0
METHOD_BATCHED = 1
METHODS = {"plain": METHOD_PLAIN, "batched": METHOD_BATCHED}


class Geometry(ctypes.Structure):
    _fields_ = [
        ("dem_width", ctypes.c_uint32),
        ("dem_size", ctypes.c_uint32),
        ("scale", ctypes.c_float),
        ("max_line_of_sight", ctypes.c_uint32),
        ("max_los_as_points", ctypes.c_uint32),
        ("computable_points_count", ctypes.c_uint32),
        ("tvs_width", ctypes.c_uint32),
        ("band_size", ctypes.c_uint32),
        ("band_deltas_size", ctypes.c_uint32),
        ("band_samples", ctypes.c_uint32),
        ("real_steps", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Params(ctypes.Structure):
    _fields_ = [
        ("dem_width", ctypes.c_uint32),
        ("scale", ctypes.c_float),
        ("max_line_of_sight", ctypes.c_uint32),
        ("rings_per_km", ctypes.c_float),
        ("observer_height", ctypes.c_float),
        ("ring_capture_sectors", ctypes.c_uint32),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("walk_ms", ctypes.c_float),
        ("reduce_ms", ctypes.c_float),
        ("bands_ms", ctypes.c_float),
        ("walk_launches", ctypes.c_uint32),
        ("point_steps", ctypes.c_uint64),
    ]


class BtHeader(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("data_size", ctypes.c_uint16),
        ("is_float", ctypes.c_uint16),
        ("horizontal_units", ctypes.c_uint16),
        ("utm_zone", ctypes.c_uint16),
        ("datum", ctypes.c_uint16),
        ("projection_source", ctypes.c_uint16),
        ("left", ctypes.c_double),
        ("right", ctypes.c_double),
        ("bottom", ctypes.c_double),
        ("top", ctypes.c_double),
        ("vertical_scale", ctypes.c_float),
    ]


class WalkShape(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_uint32),
        ("unroll", ctypes.c_uint32),
        ("xcd_contiguous", ctypes.c_uint32),
        ("pipeline", ctypes.c_uint32),
        ("persistent", ctypes.c_uint32),
        ("both_directions", ctypes.c_uint32),
        ("two_points", ctypes.c_uint32),
        ("ms", ctypes.c_float),
    ]

    def as_dict(self) -> dict:
        return {"block": self.block, "unroll": self.unroll,
                "xcd_contiguous": bool(self.xcd_contiguous),
                "directional": self.xcd_contiguous == 2,
                "pipeline": bool(self.pipeline), "persistent": bool(self.persistent),
                "both_directions": bool(self.both_directions),
                "two_points": int(self.two_points), "ms": self.ms}


class TvsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `make -C {PKG_DIR}` "
            "(or __graft_entry__.build()); the engine has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    st = ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "tvs_dem_geometry": (st, [ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, P(Geometry)]),
        "tvs_default_max_line_of_sight": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_double]),
        "tvs_ring_count_per_band": (ctypes.c_uint32, [ctypes.c_float, ctypes.c_uint32]),
        "tvs_create": (st, [P(Params), P(vp)]),
        "tvs_destroy": (st, [vp]),
        "tvs_get_geometry": (st, [vp, P(Geometry)]),
        "tvs_upload_dem_f32": (st, [vp, vp, ctypes.c_size_t]),
        "tvs_upload_dem_i16": (st, [vp, vp, ctypes.c_size_t]),
        "tvs_compute_bands": (st, [vp, ctypes.c_uint32, ctypes.c_uint32]),
        "tvs_set_band": (st, [vp, ctypes.c_uint32, vp, vp, ctypes.c_size_t]),
        "tvs_get_band": (st, [vp, ctypes.c_uint32, vp, vp, ctypes.c_size_t]),
        "tvs_compute_axes": (st, [vp, ctypes.c_uint32, vp, vp, vp]),
        "tvs_run_sectors": (st, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]),
        "tvs_run_sectors_noalloc": (st, [vp, ctypes.c_uint32, ctypes.c_uint32]),
        "tvs_observer_walk": (st, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp,
                                   vp, ctypes.c_size_t, P(ctypes.c_uint64)]),
        "tvs_run_sector_with_band": (st, [vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_size_t]),
        "tvs_reset_surfaces": (st, [vp]),
        "tvs_read_surfaces": (st, [vp, vp, ctypes.c_size_t]),
        "tvs_read_surfaces_f64": (st, [vp, vp, ctypes.c_size_t]),
        "tvs_read_rings": (st, [vp, ctypes.c_uint32, vp, ctypes.c_size_t]),
        "tvs_ring_overflow_count": (st, [vp, P(ctypes.c_uint64)]),
        "tvs_sync": (st, [vp]),
        "tvs_get_timing": (st, [vp, P(Timing)]),
        "tvs_last_error": (ctypes.c_char_p, [vp]),
        "tvs_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "tvs_bt_read_header": (st, [ctypes.c_char_p, P(BtHeader)]),
        "tvs_bt_read_f32": (st, [ctypes.c_char_p, vp, ctypes.c_size_t]),
        "tvs_bt_read_i16": (st, [ctypes.c_char_p, vp, ctypes.c_size_t]),
        "tvs_bt_write": (st, [ctypes.c_char_p, P(BtHeader), vp]),
        "tvs_bt_scale": (ctypes.c_double, [P(BtHeader)]),
        "tvs_write_heatmap_png": (st, [ctypes.c_char_p, vp, ctypes.c_uint32, ctypes.c_uint32]),
        "tvs_band_cache_save": (st, [vp, ctypes.c_char_p, ctypes.c_uint32]),
        "tvs_band_cache_load": (st, [vp, ctypes.c_char_p, ctypes.c_uint32]),
        "tvs_io_last_error": (ctypes.c_char_p, []),
        "tvs_host_alloc": (vp, [ctypes.c_size_t]),
        "tvs_host_free": (st, [vp]),
        "tvs_compact_rings": (st, [vp, ctypes.c_uint32, vp, ctypes.c_size_t, vp,
                                   ctypes.c_size_t, P(ctypes.c_uint64)]),
        "tvs_set_band_method": (st, [vp, ctypes.c_int]),
        "tvs_set_short_divide": (st, [vp, ctypes.c_uint32]),
        "tvs_set_half_heights": (st, [vp, ctypes.c_uint32]),
        "tvs_get_half_heights": (st, [vp, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]),
        "tvs_get_short_divide": (st, [vp, P(ctypes.c_uint32), P(ctypes.c_uint32)]),
        "tvs_get_fast_divide": (st, [vp, P(ctypes.c_uint32)]),
        "tvs_get_walk_history": (st, [vp, P(ctypes.c_float), ctypes.c_size_t,
                                      P(ctypes.c_size_t)]),
        "tvs_batch_set_short_divide": (st, [vp, ctypes.c_uint32]),
        "tvs_batch_set_half_heights": (st, [vp, ctypes.c_uint32]),
        "tvs_bands_sorted_count": (st, [vp, P(ctypes.c_uint32)]),
        "tvs_set_walk_shape": (st, [vp, P(WalkShape)]),
        "tvs_get_walk_shape": (st, [vp, P(WalkShape)]),
        "tvs_walk_shape_count": (ctypes.c_size_t, []),
        "tvs_autotune": (st, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                              P(WalkShape), ctypes.c_size_t, P(ctypes.c_size_t), P(WalkShape)]),
        "tvs_batch_create": (st, [P(Params), ctypes.c_uint32, P(vp)]),
        "tvs_batch_destroy": (st, [vp]),
        "tvs_batch_tile_geometry": (st, [vp, ctypes.c_uint32, P(Geometry)]),
        "tvs_batch_upload_i16": (st, [vp, ctypes.c_uint32, vp, ctypes.c_size_t]),
        "tvs_batch_upload_f32": (st, [vp, ctypes.c_uint32, vp, ctypes.c_size_t]),
        "tvs_batch_compute_bands": (st, [vp]),
        "tvs_batch_set_walk_shape": (st, [vp, P(WalkShape)]),
        "tvs_batch_get_walk_shape": (st, [vp, P(WalkShape)]),
        "tvs_batch_autotune": (st, [vp, ctypes.c_uint32, P(WalkShape), ctypes.c_size_t,
                                    P(ctypes.c_size_t), P(WalkShape)]),
        "tvs_batch_run": (st, [vp, ctypes.c_uint32]),
        "tvs_batch_sync": (st, [vp]),
        "tvs_batch_reset_surfaces": (st, [vp]),
        "tvs_batch_read_surfaces": (st, [vp, ctypes.c_uint32, vp, ctypes.c_size_t]),
        "tvs_batch_read_surfaces_f64": (st, [vp, ctypes.c_uint32, vp, ctypes.c_size_t]),
        "tvs_batch_get_timing": (st, [vp, P(Timing)]),
        "tvs_batch_last_error": (ctypes.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the engine must be contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(status: int, ctx=None) -> None:
    if status != OK:
        msg = load().tvs_last_error(ctx)
        raise TvsError(status, (msg or b"").decode() or
                       load().tvs_status_string(status).decode())


def _check_io(status: int) -> None:
    if status != OK:
        raise TvsError(status, (load().tvs_io_last_error() or b"").decode() or
                       load().tvs_status_string(status).decode())


def dem_geometry(width: int, scale: float, max_line_of_sight: int) -> Geometry:
    g = Geometry()
    _check(load().tvs_dem_geometry(width, scale, max_line_of_sight, ctypes.byref(g)))
    return g


def default_max_line_of_sight(width: int, scale: float) -> int:
    return int(load().tvs_default_max_line_of_sight(width, scale))


def ring_count_per_band(rings_per_km: float, max_line_of_sight: int) -> int:
    return int(load().tvs_ring_count_per_band(rings_per_km, max_line_of_sight))


@dataclass
class RunTiming:
    walk_ms: float
    reduce_ms: float
    bands_ms: float
    walk_launches: int
    point_steps: int


class _HostBuffer:
    """Page-locked host memory from tvs_host_alloc, freed with the last
    array that views it."""

    def __init__(self, nbytes: int):
        self.ptr = load().tvs_host_alloc(nbytes)
        if not self.ptr:
            raise MemoryError(f"tvs_host_alloc({nbytes}) failed")

    def __del__(self):
        ptr, self.ptr = getattr(self, "ptr", None), None
        if ptr and _lib is not None:       # (at interpreter exit the module may be gone)
            _lib.tvs_host_free(ptr)


def pinned_empty(n: int, dtype) -> np.ndarray:
    """An uninitialised array of n elements in page-locked host memory, so
    that device-to-host copies into it run at the link's DMA rate."""
    dtype = np.dtype(dtype)
    nbytes = max(n, 1) * dtype.itemsize
    owner = _HostBuffer(nbytes)
    raw = (ctypes.c_uint8 * nbytes).from_address(owner.ptr)
    raw._owner = owner            # the array's base keeps the allocation alive
    return np.frombuffer(raw, dtype=dtype, count=max(n, 1))


class Engine:
    """One device context (tvs_ctx): a DEM, its resident band data, its
    cumulative surfaces and optional ring records."""

    def __init__(self, dem_width: int, scale: float, max_line_of_sight: int = 0,
                 rings_per_km: float = 5.0, observer_height: float = 1.65,
                 ring_capture_sectors: int = 0):
        L = load()
        p = Params(dem_width, scale, max_line_of_sight, rings_per_km, observer_height,
                   ring_capture_sectors)
        h = ctypes.c_void_p()
        _check(L.tvs_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.geometry = Geometry()
        _check(L.tvs_get_geometry(self._h, ctypes.byref(self.geometry)), self._h)
        self.rings_per_band = ring_count_per_band(rings_per_km,
                                                  self.geometry.max_line_of_sight)
        self.observer_height = observer_height

    # -- lifetime --
    def close(self) -> None:
        if self._h:
            load().tvs_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, status: int) -> None:
        _check(status, self._h)

    # -- inputs --
    @property
    def n_points(self) -> int:
        return self.geometry.computable_points_count

    def upload(self, elevations: np.ndarray) -> None:
        """int16 arrays go up as int16 (the batched method reads them at 2 B
        per step); anything else as float32 (which the library stores as
        int16 when every height is an int16 value)."""
        if elevations.dtype == np.int16:
            a = np.ascontiguousarray(elevations.reshape(-1))
            self._c(load().tvs_upload_dem_i16(self._h, _ptr(a), a.size))
        else:
            a = np.ascontiguousarray(elevations.reshape(-1), np.float32)
            self._c(load().tvs_upload_dem_f32(self._h, _ptr(a), a.size))

    def compute_bands(self, first: int = 0, count: int = 180) -> None:
        self._c(load().tvs_compute_bands(self._h, first, count))

    def set_band_method(self, method: str) -> None:
        """'select' (default: exact window selection, no full sorts) or
        'sort' (the two stable radix sorts of Axes::compute)."""
        self._c(load().tvs_set_band_method(self._h, {"select": 0, "sort": 1}[method]))

    def bands_sorted_count(self) -> int:
        v = ctypes.c_uint32(0)
        self._c(load().tvs_bands_sorted_count(self._h, ctypes.byref(v)))
        return int(v.value)

    def set_band(self, sector: int, deltas: np.ndarray, dists: np.ndarray) -> None:
        d = np.ascontiguousarray(deltas, np.int32)
        s = np.ascontiguousarray(dists, np.float32)
        self._c(load().tvs_set_band(self._h, sector, _ptr(d), _ptr(s), d.size))

    def get_band(self, sector: int):
        n = self.geometry.band_deltas_size
        d = np.empty(max(n, 1), np.int32)
        s = np.empty(max(n, 1), np.float32)
        self._c(load().tvs_get_band(self._h, sector, _ptr(d), _ptr(s), n))
        return d[:n], s[:n]

    def compute_axes(self, sector: int):
        n = self.geometry.dem_size
        dist = np.empty(n, np.float32)
        sight = np.empty(n, np.uint32)
        order = np.empty(n, np.uint32)
        self._c(load().tvs_compute_axes(self._h, sector, _ptr(dist), _ptr(sight), _ptr(order)))
        return dist, sight, order

    # -- walk --
    def run(self, first: int = 0, count: int = 180, method: str = "batched") -> None:
        self._c(load().tvs_run_sectors(self._h, first, count, METHODS[method]))

    def run_noalloc(self, first: int = 0, count: int = 180) -> None:
        """tvs_run_sectors_noalloc: the batched run with no device allocation
        (band surfaces in the library's static scratch); synchronous."""
        self._c(load().tvs_run_sectors_noalloc(self._h, first, count))

    def run_sector_with_band(self, deltas: np.ndarray, dists: np.ndarray) -> np.ndarray:
        d = np.ascontiguousarray(deltas, np.int32)
        s = np.ascontiguousarray(dists, np.float32)
        out = np.empty(max(self.n_points, 1), np.float32)
        self._c(load().tvs_run_sector_with_band(self._h, _ptr(d), _ptr(s), d.size,
                                                _ptr(out), self.n_points))
        return out[:self.n_points]

    def reset_surfaces(self) -> None:
        self._c(load().tvs_reset_surfaces(self._h))

    def surfaces(self) -> np.ndarray:
        out = np.empty(max(self.n_points, 1), np.float32)
        self._c(load().tvs_read_surfaces(self._h, _ptr(out), self.n_points))
        return out[:self.n_points]

    def surfaces_f64(self) -> np.ndarray:
        out = np.empty(max(self.n_points, 1), np.float64)
        self._c(load().tvs_read_surfaces_f64(self._h, _ptr(out), self.n_points))
        return out[:self.n_points]

    def rings(self, slot: int) -> np.ndarray:
        n = 2 * self.n_points * self.rings_per_band
        out = np.empty(max(n, 1), np.uint32)
        self._c(load().tvs_read_rings(self._h, slot, _ptr(out), n))
        return out[:n]

    def compact_rings(self, sector: int, pinned: bool = True):
        """(counts u32[2N], events u32[total]) of one sector: counts[k] is the
        count slot of band k, events its ids after the PoV slot.  With
        `pinned` both arrays live in page-locked host memory
        (tvs_host_alloc), which the device writes at DMA rate; where that
        memory cannot be had (a locked-memory limit), ordinary arrays."""
        def empty(n, t):
            if pinned:
                try:
                    return pinned_empty(n, t)
                except MemoryError:
                    pass
            return np.empty(max(n, 1), t)
        bands = 2 * self.n_points
        counts = empty(bands, np.uint32)
        total = ctypes.c_uint64(0)
        self._c(load().tvs_compact_rings(self._h, sector, _ptr(counts), bands, None, 0,
                                         ctypes.byref(total)))
        events = empty(int(total.value), np.uint32)
        self._c(load().tvs_compact_rings(self._h, sector, _ptr(counts), bands, _ptr(events),
                                         events.size, ctypes.byref(total)))
        return counts[:bands], events[:int(total.value)]

    def observer_walk(self, tvs_id: int, first: int = 0, count: int = 180):
        """tvs_observer_walk: one observer's bands (sector after sector,
        forward then backward), one wavefront per band.  Returns (counts
        u32[2*count], band surfaces f32[2*count], events u32[total])."""
        bands = 2 * count
        counts = np.empty(max(bands, 1), np.uint32)
        surf = np.empty(max(bands, 1), np.float32)
        total = ctypes.c_uint64(0)
        self._c(load().tvs_observer_walk(self._h, tvs_id, first, count, _ptr(counts), _ptr(surf),
                                         None, 0, ctypes.byref(total)))
        events = np.empty(max(int(total.value), 1), np.uint32)
        self._c(load().tvs_observer_walk(self._h, tvs_id, first, count, _ptr(counts), _ptr(surf),
                                         _ptr(events), events.size, ctypes.byref(total)))
        return counts[:bands], surf[:bands], events[:int(total.value)]

    def ring_overflow(self) -> int:
        v = ctypes.c_uint64(0)
        self._c(load().tvs_ring_overflow_count(self._h, ctypes.byref(v)))
        return int(v.value)

    def set_walk_shape(self, block: int, unroll: int, xcd_contiguous: bool = True,
                       pipeline: bool = False, persistent: bool = False,
                       both_directions: bool = False, two_points: int = 0) -> None:
        self._c(load().tvs_set_walk_shape(self._h, ctypes.byref(
            WalkShape(block, unroll, int(xcd_contiguous), int(pipeline), int(persistent),
                      int(both_directions), int(two_points), 0.0))))

    def walk_shape(self) -> dict:
        w = WalkShape()
        self._c(load().tvs_get_walk_shape(self._h, ctypes.byref(w)))
        return w.as_dict()

    def set_short_divide(self, enable: bool) -> None:
        self._c(load().tvs_set_short_divide(self._h, int(enable)))

    def set_half_heights(self, enable: bool) -> None:
        self._c(load().tvs_set_half_heights(self._h, int(enable)))

    def half_heights(self) -> tuple[bool, bool]:
        """(enabled, the uploaded DEM has the f16 form)."""
        en, av = ctypes.c_uint32(0), ctypes.c_uint32(0)
        self._c(load().tvs_get_half_heights(self._h, ctypes.byref(en), ctypes.byref(av)))
        return bool(en.value), bool(av.value)

    def short_divide(self) -> tuple[bool, int]:
        """(enabled, sectors verified for the two-operation divide)."""
        en, n = ctypes.c_uint32(0), ctypes.c_uint32(0)
        self._c(load().tvs_get_short_divide(self._h, ctypes.byref(en), ctypes.byref(n)))
        return bool(en.value), int(n.value)

    def walk_history(self, n: int) -> np.ndarray:
        """Walk times (ms) of the last n runs (at most 64), oldest first."""
        out = (ctypes.c_float * max(n, 1))()
        got = ctypes.c_size_t(0)
        self._c(load().tvs_get_walk_history(self._h, out, n, ctypes.byref(got)))
        return np.array(out[:got.value], np.float64)

    def fast_divide(self) -> bool:
        """Whether the uploaded heights admit the four-operation divide
        (else the batched walk uses the IEEE divide)."""
        v = ctypes.c_uint32(0)
        self._c(load().tvs_get_fast_divide(self._h, ctypes.byref(v)))
        return bool(v.value)

    def autotune(self, first: int = 0, count: int = 180, reps: int = 2):
        """Time every compiled launch shape on this DEM, keep the fastest."""
        n = int(load().tvs_walk_shape_count())
        res = (WalkShape * n)()
        got = ctypes.c_size_t(0)
        best = WalkShape()
        self._c(load().tvs_autotune(self._h, first, count, reps, res, n, ctypes.byref(got),
                                    ctypes.byref(best)))
        return best.as_dict(), [res[i].as_dict() for i in range(got.value)]

    def cache_save(self, state_dir: str, sector: int) -> None:
        self._c(load().tvs_band_cache_save(self._h, state_dir.encode(), sector))

    def cache_load(self, state_dir: str, sector: int) -> None:
        self._c(load().tvs_band_cache_load(self._h, state_dir.encode(), sector))

    def sync(self) -> None:
        self._c(load().tvs_sync(self._h))

    def timing(self) -> RunTiming:
        t = Timing()
        self._c(load().tvs_get_timing(self._h, ctypes.byref(t)))
        return RunTiming(t.walk_ms, t.reduce_ms, t.bands_ms, t.walk_launches, t.point_steps)


class Batch:
    """Many DEM tiles whose 180 sectors run in one launch (tvs_batch)."""

    def __init__(self, tiles):
        """tiles: iterable of (dem_width, scale, max_line_of_sight,
        rings_per_km, observer_height)."""
        tiles = list(tiles)
        arr = (Params * len(tiles))(*[Params(w, s, los, rpk, obs, 0)
                                      for (w, s, los, rpk, obs) in tiles])
        h = ctypes.c_void_p()
        st = load().tvs_batch_create(arr, len(tiles), ctypes.byref(h))
        if st != OK:
            raise TvsError(st, (load().tvs_batch_last_error(None) or b"").decode())
        self._h = h
        self.n_tiles = len(tiles)
        self.geometry = []
        for i in range(self.n_tiles):
            g = Geometry()
            self._c(load().tvs_batch_tile_geometry(self._h, i, ctypes.byref(g)))
            self.geometry.append(g)

    def _c(self, status: int) -> None:
        if status != OK:
            raise TvsError(status, (load().tvs_batch_last_error(self._h) or b"").decode())

    def close(self) -> None:
        if self._h:
            load().tvs_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, tile: int, elevations: np.ndarray) -> None:
        if elevations.dtype == np.int16:
            a = np.ascontiguousarray(elevations.reshape(-1))
            self._c(load().tvs_batch_upload_i16(self._h, tile, _ptr(a), a.size))
        else:
            a = np.ascontiguousarray(elevations.reshape(-1), np.float32)
            self._c(load().tvs_batch_upload_f32(self._h, tile, _ptr(a), a.size))

    def compute_bands(self) -> None:
        self._c(load().tvs_batch_compute_bands(self._h))

    def set_walk_shape(self, block: int, unroll: int, xcd_contiguous: bool = False,
                       pipeline: bool = False, persistent: bool = False,
                       both_directions: bool = False, two_points: int = 0) -> None:
        self._c(load().tvs_batch_set_walk_shape(self._h, ctypes.byref(
            WalkShape(block, unroll, int(xcd_contiguous), int(pipeline), int(persistent),
                      int(both_directions), int(two_points), 0.0))))

    def walk_shape(self) -> dict:
        w = WalkShape()
        self._c(load().tvs_batch_get_walk_shape(self._h, ctypes.byref(w)))
        return w.as_dict()

    def set_short_divide(self, enable: bool) -> None:
        self._c(load().tvs_batch_set_short_divide(self._h, int(enable)))

    def set_half_heights(self, enable: bool) -> None:
        """Engine.set_half_heights for every tile."""
        self._c(load().tvs_batch_set_half_heights(self._h, int(enable)))

    def autotune(self, reps: int = 1):
        """Time every compiled launch shape on the whole batch, keep the
        fastest; cumulative surfaces are unchanged."""
        n = int(load().tvs_walk_shape_count())
        res = (WalkShape * n)()
        got = ctypes.c_size_t(0)
        best = WalkShape()
        self._c(load().tvs_batch_autotune(self._h, reps, res, n, ctypes.byref(got),
                                          ctypes.byref(best)))
        return best.as_dict(), [res[i].as_dict() for i in range(got.value)]

    def run(self, use_graph: bool = True) -> None:
        self._c(load().tvs_batch_run(self._h, int(use_graph)))

    def sync(self) -> None:
        self._c(load().tvs_batch_sync(self._h))

    def reset_surfaces(self) -> None:
        self._c(load().tvs_batch_reset_surfaces(self._h))

    def surfaces(self, tile: int) -> np.ndarray:
        n = self.geometry[tile].computable_points_count
        out = np.empty(max(n, 1), np.float32)
        self._c(load().tvs_batch_read_surfaces(self._h, tile, _ptr(out), n))
        return out[:n]

    def surfaces_f64(self, tile: int) -> np.ndarray:
        n = self.geometry[tile].computable_points_count
        out = np.empty(max(n, 1), np.float64)
        self._c(load().tvs_batch_read_surfaces_f64(self._h, tile, _ptr(out), n))
        return out[:n]

    def timing(self) -> RunTiming:
        t = Timing()
        self._c(load().tvs_batch_get_timing(self._h, ctypes.byref(t)))
        return RunTiming(t.walk_ms, t.reduce_ms, t.bands_ms, t.walk_launches, t.point_steps)
