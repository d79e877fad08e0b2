"""Host-side mirror of the reference's `DEM` and `Compute` interface
(crates/total-viewsheds/src/dem.rs:11-139, compute.rs:6-115) over the C ABI.

`Compute(method, state_directory, output_dir, dem, rings_per_km).run()`
returns `(cumulative_surfaces, ring_data)` as compute.rs:101-115 does; the
backend slot that compute.rs:197 leaves `unimplemented!()` is the HIP engine
("hip" = batched method, "hip-plain" = one sector per launch).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib, bt

SECTOR_STEPS = 180  # axes.rs:35


class DEM:
    """dem.rs:11-84: geometry of a square DEM plus its elevations."""

    def __init__(self, width: int, scale: float, max_line_of_sight: int):
        g = _lib.dem_geometry(width, scale, max_line_of_sight)
        self.width = g.dem_width
        self.size = g.dem_size
        self.scale = g.scale
        self.max_line_of_sight = g.max_line_of_sight
        self.max_los_as_points = g.max_los_as_points
        self.computable_points_count = g.computable_points_count
        self.tvs_width = g.tvs_width
        self.band_size = g.band_size
        self.geometry = g
        self.elevations: np.ndarray | None = None

    def band_deltas_size(self) -> int:
        """band_of_sight.rs:54-58"""
        return self.band_size - 1

    def tvs_id_to_pov_id(self, tvs_id: int) -> int:
        """dem.rs:127-132"""
        x = tvs_id % self.tvs_width + self.max_los_as_points
        y = tvs_id // self.tvs_width + self.max_los_as_points
        return y * self.width + x

    def pov_id_to_tvs_id(self, pov_id: int) -> int:
        """dem.rs:118-124"""
        x = pov_id % self.width - self.max_los_as_points
        y = pov_id // self.width - self.max_los_as_points
        return y * self.tvs_width + x


class Compute:
    """compute.rs:6-115 with the HIP engine as the backend.

    `state_directory` (or None) keeps the band-data cache of compute.rs:118-161
    in the reference's layout (`<state>/total-viewsheds/<width>/{band_deltas,
    distances}/<angle>.bin`, cache.rs:23-76); a cached sector is loaded
    instead of computed, a missing one is computed on the device and saved.
    A cache file of another band length (cache.rs keys on the width only) is
    recomputed and overwritten instead of being used.  `output_dir` (or None)
    receives `heatmap.png` after every sector, rendered from the running
    total of the cumulative surfaces (compute.rs:165-178, 206-214)."""

    def __init__(self, method: str, state_directory, output_dir, dem: DEM,
                 rings_per_km: float, observer_height: float = 1.8,
                 capture_rings: bool = True):
        if method not in ("hip", "hip-plain"):
            raise ValueError(f"compute method {method!r}: this engine provides 'hip' "
                             "(batched) and 'hip-plain' (one sector per launch)")
        if dem.elevations is None:
            raise ValueError("DEM has no elevations")
        self.method = "batched" if method == "hip" else "plain"
        self.state_directory = None if state_directory is None else os.fspath(state_directory)
        self.output_dir = None if output_dir is None else os.fspath(output_dir)
        self.dem = dem
        self.rings_per_km = rings_per_km
        self.capture_rings = capture_rings
        self.engine = _lib.Engine(dem.width, dem.scale, dem.max_line_of_sight,
                                  rings_per_km, observer_height,
                                  SECTOR_STEPS if capture_rings else 0)
        self.engine.upload(dem.elevations)
        self.reserved_rings_per_band = self.engine.rings_per_band
        self.total_surfaces = np.zeros(dem.computable_points_count, np.float32)
        self.sectors_loaded = 0     # sectors whose band data came from the cache

    def run(self):
        """compute.rs:101-115: all 180 sectors; returns the cumulative
        surfaces and, per sector, the ring records (test-mode layout)."""
        e = self.engine
        e.reset_surfaces()
        self.total_surfaces[:] = 0.0
        self.sectors_loaded = 0
        if self.state_directory is None and self.output_dir is None:
            # Nothing happens between sectors: one call for all of them.
            e.compute_bands(0, SECTOR_STEPS)
            e.run(0, SECTOR_STEPS, self.method)
            surfaces = e.surfaces()
            rings = [e.rings(s) for s in range(SECTOR_STEPS)] if self.capture_rings else []
            return surfaces, rings
        rings = []
        for angle in range(SECTOR_STEPS):
            self.load_or_compute_cache(angle)
            e.run(angle, 1, self.method)
            if self.capture_rings:
                rings.append(e.rings(0))
            cumulative = e.surfaces()
            self.add_sector_surfaces_to_running_total(cumulative)
            self.render_total_surfaces()
        return e.surfaces(), rings

    def load_or_compute_cache(self, angle: int) -> None:
        """compute.rs:118-161."""
        e = self.engine
        if self.state_directory is not None:
            try:
                e.cache_load(self.state_directory, angle)
                self.sectors_loaded += 1
                return
            except _lib.TvsError as err:
                if err.status not in (_lib.ERR_IO, _lib.ERR_SIZE):
                    raise
        e.compute_bands(angle, 1)
        if self.state_directory is not None:
            e.cache_save(self.state_directory, angle)

    def add_sector_surfaces_to_running_total(self, cumulative: np.ndarray) -> None:
        """compute.rs:206-214 (f32 adds)."""
        self.total_surfaces += cumulative

    def render_total_surfaces(self) -> None:
        """compute.rs:165-178: heatmap.png of the running total."""
        if self.output_dir is None:
            return
        os.makedirs(self.output_dir, exist_ok=True)
        T = self.dem.tvs_width
        bt.heatmap_png(os.path.join(self.output_dir, "heatmap.png"),
                       self.total_surfaces.reshape(T, T))
