"""ctypes view of the oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (``tvs_oracle.c``) is the CPU restatement of the reference
total-viewshed path that the HIP engine is checked against.  Only the tests,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this module; the engine never does.

``output_ascii`` restates the reference's test-only ring decoder
(crates/total-viewsheds/src/output/ascii.rs:24-104) so the reference's
ASCII viewshed fixtures can be checked.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TVS_ORACLE_LIB names another build of the same source (the sanitized one,
# or a variant with a planted mistake that a test expects to be caught).
LIB_PATH = os.environ.get("TVS_ORACLE_LIB") or os.path.join(HERE, "libtvs_oracle.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)


class _Dem(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("size", ctypes.c_uint32),
        ("scale", ctypes.c_float),
        ("max_line_of_sight", ctypes.c_uint32),
        ("max_los_as_points", ctypes.c_uint32),
        ("computable_points_count", ctypes.c_uint32),
        ("tvs_width", ctypes.c_uint32),
        ("band_size", ctypes.c_uint32),
    ]


class _Constants(ctypes.Structure):
    _fields_ = [
        ("total_bands", ctypes.c_uint32),
        ("max_los_as_points", ctypes.c_uint32),
        ("dem_width", ctypes.c_uint32),
        ("tvs_width", ctypes.c_uint32),
        ("observer_height", ctypes.c_float),
        ("reserved_rings_per_band", ctypes.c_uint32),
    ]


def build() -> None:
    """Compile the oracle in place (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH) and "TVS_ORACLE_LIB" not in os.environ:
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.tvso_dem_new.argtypes = [ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32,
                                   ctypes.POINTER(_Dem)]
        L.tvso_default_max_line_of_sight.argtypes = [ctypes.c_uint32, ctypes.c_double]
        L.tvso_default_max_line_of_sight.restype = ctypes.c_uint32
        L.tvso_ring_count_per_band.argtypes = [ctypes.c_float, ctypes.c_uint32]
        L.tvso_ring_count_per_band.restype = ctypes.c_uint32
        L.tvso_shifted_angle.argtypes = [ctypes.c_uint16]
        L.tvso_shifted_angle.restype = ctypes.c_float
        L.tvso_calculate_distances.argtypes = [ctypes.c_uint32, ctypes.c_float, vp]
        L.tvso_order_by_distance.argtypes = [vp, ctypes.c_size_t, vp]
        L.tvso_axes_compute.argtypes = [ctypes.c_uint32, ctypes.c_uint16, vp, vp, vp]
        L.tvso_compile_band_data.argtypes = [ctypes.POINTER(_Dem), vp, vp, vp, vp, vp]
        L.tvso_band_data.argtypes = [ctypes.POINTER(_Dem), ctypes.c_uint16, vp, vp]
        L.tvso_kernel.argtypes = [ctypes.c_uint32, ctypes.POINTER(_Constants), vp,
                                  ctypes.c_size_t, vp, vp, ctypes.c_uint32, vp, vp,
                                  ctypes.c_size_t, vp]
        L.tvso_compute_run.argtypes = [ctypes.POINTER(_Dem), vp, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_uint16, ctypes.c_uint16,
                                       vp, vp, vp, ctypes.c_int]
        L.tvso_walk_sectors.argtypes = [ctypes.POINTER(_Dem), vp, ctypes.c_float,
                                        ctypes.c_uint32, vp, vp, ctypes.c_uint16, vp,
                                        vp, ctypes.c_int]
        L.tvso_walk_sectors_f64.argtypes = [ctypes.POINTER(_Dem), vp, ctypes.c_float,
                                            vp, vp, ctypes.c_uint16, vp, vp, vp,
                                            ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


def _check(st: int, what: str) -> None:
    if st != 0:
        raise OracleError(f"{what}: oracle status {st}")


@dataclass
class Dem:
    """dem.rs:11-84 scalars."""
    width: int
    size: int
    scale: float
    max_line_of_sight: int
    max_los_as_points: int
    computable_points_count: int
    tvs_width: int
    band_size: int

    @property
    def n_deltas(self) -> int:
        return self.band_size - 1

    def c(self) -> _Dem:
        return _Dem(self.width, self.size, self.scale, self.max_line_of_sight,
                    self.max_los_as_points, self.computable_points_count,
                    self.tvs_width, self.band_size)

    def tvs_id_to_pov_id(self, tvs_id: int) -> int:
        """dem.rs:127-132"""
        x = tvs_id % self.tvs_width + self.max_los_as_points
        y = tvs_id // self.tvs_width + self.max_los_as_points
        return y * self.width + x


def dem_new(width: int, scale: float, max_line_of_sight: int) -> Dem:
    d = _Dem()
    _check(lib().tvso_dem_new(width, scale, max_line_of_sight, ctypes.byref(d)),
           "DEM::new")
    return Dem(d.width, d.size, d.scale, d.max_line_of_sight, d.max_los_as_points,
               d.computable_points_count, d.tvs_width, d.band_size)


def default_max_line_of_sight(width: int, scale: float) -> int:
    return int(lib().tvso_default_max_line_of_sight(width, scale))


def ring_count_per_band(rings_per_km: float, max_line_of_sight: int) -> int:
    return int(lib().tvso_ring_count_per_band(rings_per_km, max_line_of_sight))


def shifted_angle(sector: int) -> float:
    return float(lib().tvso_shifted_angle(sector))


def calculate_distances(width: int, angle: float) -> np.ndarray:
    out = np.empty(width * width, np.float64)
    lib().tvso_calculate_distances(width, angle, _ptr(out))
    return out


def order_by_distance(distances: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(distances, np.float64)
    out = np.empty(d.size, np.uint32)
    _check(lib().tvso_order_by_distance(_ptr(d), d.size, _ptr(out)), "order_by_distance")
    return out


def axes_compute(width: int, sector: int):
    """axes.rs:103-130 for Axes::new(width, sector): (distances f32,
    sight_ordered_map, sector_ordered)."""
    n = width * width
    dist = np.empty(n, np.float32)
    sight = np.empty(n, np.uint32)
    sector_ordered = np.empty(n, np.uint32)
    _check(lib().tvso_axes_compute(width, sector, _ptr(dist), _ptr(sight),
                                   _ptr(sector_ordered)), "Axes::compute")
    return dist, sight, sector_ordered


def compile_band_data(dem: Dem, distances, sight_map, sector_ordered):
    deltas = np.empty(max(dem.n_deltas, 1), np.int32)
    dists = np.empty(max(dem.n_deltas, 1), np.float32)
    d = dem.c()
    _check(lib().tvso_compile_band_data(ctypes.byref(d), _ptr(distances), _ptr(sight_map),
                                        _ptr(sector_ordered), _ptr(deltas), _ptr(dists)),
           "compile_band_data")
    return deltas[:dem.n_deltas], dists[:dem.n_deltas]


def band_data(dem: Dem, sector: int):
    deltas = np.empty(max(dem.n_deltas, 1), np.int32)
    dists = np.empty(max(dem.n_deltas, 1), np.float32)
    d = dem.c()
    _check(lib().tvso_band_data(ctypes.byref(d), sector, _ptr(deltas), _ptr(dists)),
           f"band data (sector {sector})")
    return deltas[:dem.n_deltas], dists[:dem.n_deltas]


def band_data_all(dem: Dem, sectors) -> tuple[np.ndarray, np.ndarray]:
    sectors = list(sectors)
    deltas = np.zeros((len(sectors), dem.n_deltas), np.int32)
    dists = np.zeros((len(sectors), dem.n_deltas), np.float32)
    for i, s in enumerate(sectors):
        deltas[i], dists[i] = band_data(dem, s)
    return deltas, dists


def compute_run(dem: Dem, elevations: np.ndarray, observer_height: float = 1.8,
                rings_per_km: float = 5000.0, first_sector: int = 0,
                sector_count: int = 180, rings: bool = True, n_threads: int = 1,
                cumulative: np.ndarray | None = None):
    """compute.rs:101-115 (CPU backend, test-mode ring sizing).

    Returns (cumulative_surfaces f32[N], ring_data u32[sectors, 2N*R] or None,
    ring_spill)."""
    elev = np.ascontiguousarray(elevations, np.float32)
    assert elev.size == dem.size
    N = dem.computable_points_count
    cum = np.zeros(N, np.float32) if cumulative is None else cumulative
    R = ring_count_per_band(rings_per_km, dem.max_line_of_sight)
    ring = np.zeros((sector_count, 2 * N * R), np.uint32) if rings else None
    spill = ctypes.c_uint64(0)
    d = dem.c()
    _check(lib().tvso_compute_run(ctypes.byref(d), _ptr(elev), observer_height,
                                  rings_per_km, first_sector, sector_count, _ptr(cum),
                                  _ptr(ring), ctypes.byref(spill), n_threads),
           "Compute::run")
    return cum, ring, int(spill.value)


def kernel_band(dem: Dem, elevations: np.ndarray, observer_height: float,
                deltas: np.ndarray, dists: np.ndarray, kernel_id: int):
    """kernel.rs:71-256 for one kernel id over one sector's band data: the
    band surface (added to a zero cumulative surface, so exactly the band's)
    and its ring record [count, PoV, events...] with room for every event."""
    elev = np.ascontiguousarray(elevations, np.float32)
    deltas = np.ascontiguousarray(deltas, np.int32)
    dists = np.ascontiguousarray(dists, np.float32)
    n = deltas.size
    R = n + 3
    c = _Constants(2 * dem.computable_points_count, dem.max_los_as_points, dem.width,
                   dem.tvs_width, observer_height, R)
    cum = np.zeros(dem.computable_points_count, np.float32)
    ring = np.zeros((2 * dem.computable_points_count) * R, np.uint32)
    _check(lib().tvso_kernel(kernel_id, ctypes.byref(c), _ptr(elev), elev.size, _ptr(dists),
                             _ptr(deltas), n, _ptr(cum), _ptr(ring), ring.size, None),
           f"kernel {kernel_id}")
    t = kernel_id % dem.computable_points_count
    rec = ring[kernel_id * R:(kernel_id + 1) * R]
    return cum[t], rec[:int(rec[0]) + 1]


def walk_sectors(dem: Dem, elevations: np.ndarray, observer_height: float,
                 deltas: np.ndarray, dists: np.ndarray, rings_per_band: int = 0,
                 rings: bool = False, n_threads: int = 1,
                 cumulative: np.ndarray | None = None):
    """The walk alone (compute.rs:228-240 per sector) over given band data."""
    elev = np.ascontiguousarray(elevations, np.float32)
    deltas = np.ascontiguousarray(deltas, np.int32)
    dists = np.ascontiguousarray(dists, np.float32)
    S = deltas.shape[0]
    N = dem.computable_points_count
    cum = np.zeros(N, np.float32) if cumulative is None else cumulative
    ring = np.zeros((S, 2 * N * rings_per_band), np.uint32) if rings else None
    d = dem.c()
    _check(lib().tvso_walk_sectors(ctypes.byref(d), _ptr(elev), observer_height,
                                   rings_per_band, _ptr(deltas), _ptr(dists), S,
                                   _ptr(cum), _ptr(ring), n_threads), "walk")
    return cum, ring


def walk_sectors_f64(dem: Dem, elevations: np.ndarray, observer_height: float,
                     deltas: np.ndarray, dists: np.ndarray, n_threads: int = 1):
    """FP64 walk over the same band data: (tvs64, flips, steps)."""
    elev = np.ascontiguousarray(elevations, np.float32)
    deltas = np.ascontiguousarray(deltas, np.int32)
    dists = np.ascontiguousarray(dists, np.float32)
    S = deltas.shape[0]
    tvs64 = np.zeros(dem.computable_points_count, np.float64)
    flips = ctypes.c_uint64(0)
    steps = ctypes.c_uint64(0)
    d = dem.c()
    _check(lib().tvso_walk_sectors_f64(ctypes.byref(d), _ptr(elev), observer_height,
                                       _ptr(deltas), _ptr(dists), S, _ptr(tvs64),
                                       ctypes.byref(flips), ctypes.byref(steps),
                                       n_threads), "walk f64")
    return tvs64, int(flips.value), int(steps.value)


def output_ascii(dem: Dem, pov_id: int, all_ring_data, reserved_ring_size: int):
    """output/ascii.rs:24-104: rebuild one observer's viewshed as ASCII rows
    ('+' opening, '-' closing, '±' both, 'o' the observer)."""
    view = ["."] * dem.size
    for sector in all_ring_data:
        it = iter(int(v) for v in sector)
        for _ in range(dem.computable_points_count * 2):        # ascii.rs:61
            no_of_ring_values = next(it) // 2                   # :63
            band_pov = next(it)                                 # :66
            for index in range(no_of_ring_values):              # :68-79
                opening = band_pov if index == 0 else next(it)
                closing = next(it)
                if band_pov == pov_id:
                    _populate(view, opening, closing)
            skip = reserved_ring_size - no_of_ring_values * 2 - 2   # :81
            for _ in range(skip + 1):                           # Iterator::nth(skip)
                next(it, None)
    view[pov_id] = "o"
    w = dem.width
    return [" ".join(view[r * w:(r + 1) * w]) for r in range(w)]


def _populate(view, opening: int, closing: int) -> None:
    """output/ascii.rs:88-104"""
    view[opening] = "+" if view[opening] == "." else "±"
    view[closing] = "-" if view[closing] == "." else "±"
    if opening == closing:
        view[opening] = "±"
