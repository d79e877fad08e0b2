"""MI355X total-viewshed engine: Python view of the C ABI (include/tvs.h)."""
from ._lib import (Batch, Engine, Geometry, TvsError, default_max_line_of_sight,  # noqa: F401
                   dem_geometry, load, ring_count_per_band, LIB_PATH, HEADER_PATH)
from .compute import DEM, Compute, SECTOR_STEPS  # noqa: F401
