// geometry.cpp -- host-side DEM geometry: the scalars of `DEM::new`
// (dem.rs:44-84), the band template sizes (band_of_sight.rs:54-58, 95,
// 149-150) and the per-sector rotation constants (axes.rs:89-100, 154-158).
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace tvs {

namespace {

// Rust `v as u32` for f32/f64: saturating, NaN -> 0.
template <typename F>
uint32_t to_u32_saturating(F v) {
    if (!(v > F(0))) return 0u;
    if (v >= F(4294967296.0)) return 0xffffffffu;
    return static_cast<uint32_t>(v);
}

uint32_t isqrt(uint64_t n) {
    uint64_t r = static_cast<uint64_t>(std::sqrt(static_cast<double>(n)));
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return static_cast<uint32_t>(r);
}

}  // namespace

// dem.rs:44-84.  The computable test (dem.rs:106-114) is the same f32
// predicate on x and on y, so the count is (computable columns)^2.
tvs_status dem_geometry(uint32_t width, float scale, uint32_t los,
                        tvs_geometry *out) {
    if (!out || width == 0 || !(scale > 0.0f)) return TVS_ERR_INVALID_ARGUMENT;
    if (width > 46340u) return TVS_ERR_INVALID_ARGUMENT;  // w*w fits in u32
    tvs_geometry g;
    std::memset(&g, 0, sizeof g);
    g.dem_width = width;
    g.dem_size = width * width;
    g.scale = scale;
    g.max_line_of_sight = los;
    g.max_los_as_points = to_u32_saturating(static_cast<float>(los) / scale);
    if (g.max_los_as_points > width / 2u) return TVS_ERR_LOS_TOO_LONG;
    const float lower = static_cast<float>(los);
    const float upper = static_cast<float>(width - 1u) * scale - lower;
    uint64_t columns = 0;
    for (uint32_t x = 0; x < width; ++x) {
        const float fx = static_cast<float>(x) * scale;
        if (fx >= lower && fx <= upper) ++columns;
    }
    g.computable_points_count = static_cast<uint32_t>(columns * columns);
    g.tvs_width = isqrt(g.computable_points_count);
    g.band_size = g.max_los_as_points + 1u;
    g.band_deltas_size = g.band_size - 1u;
    g.band_samples = static_cast<uint32_t>(
        static_cast<double>(g.band_size * 2u) * std::sqrt(2.0));
    const uint32_t middle = g.band_samples / 2u;
    const uint32_t filled = 2u * g.band_size > middle ? 2u * g.band_size - middle : 0u;
    g.real_steps = filled < g.band_deltas_size ? filled : g.band_deltas_size;
    *out = g;
    return TVS_OK;
}

// The sine and cosine that axes.rs:154-158 feed into each sector's
// rotated-coordinate distances: the sight axis at `sector + 0.001` degrees
// (f32), the sector axis 90 degrees on (f32 add, axes.rs:127).  Host libm,
// as the reference's f64::sin/cos.
void sector_sin_cos(uint32_t sector, bool sector_axis, double *s, double *c) {
    float angle = static_cast<float>(sector) + kSectorShift;
    if (sector_axis) angle = angle + 90.0f;
    const double rad = static_cast<double>(angle) * (M_PI / 180.0);
    *s = std::sin(rad);
    *c = std::cos(rad);
}

}  // namespace tvs

extern "C" {

tvs_status tvs_dem_geometry(uint32_t dem_width, float scale,
                            uint32_t max_line_of_sight, tvs_geometry *out) {
    return tvs::dem_geometry(dem_width, scale, max_line_of_sight, out);
}

uint32_t tvs_default_max_line_of_sight(uint32_t dem_width, double scale) {
    return tvs::to_u32_saturating(static_cast<double>(dem_width / 3u) * scale);
}

uint32_t tvs_ring_count_per_band(float rings_per_km,
                                 uint32_t max_line_of_sight) {
    const float km = static_cast<float>(max_line_of_sight) / 1000.0f;
    return tvs::to_u32_saturating(km * rings_per_km);
}

}  // extern "C"
