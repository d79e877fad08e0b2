// tune.cpp -- launch-shape selection for the batched walk.
//
// The shapes compiled into the library (points per work item = threads per
// block, steps unrolled per chunk, software pipelining, XCD-contiguous or
// round-robin item mapping, a persistent grid, both bands of a point per
// lane, two points per lane) are timed on the context's own DEM and band data; the fastest
// becomes the context's shape.  The cumulative surfaces are saved and
// restored around the timing runs; the timing runs record no rings and leave
// the ring records and overflow count of the caller's last run in place.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"

namespace tvs {

std::vector<WalkConfig> compiled_walk_shapes() {
    std::vector<WalkConfig> out;
    // A persistent grid strides over the items itself: one item mapping.
    // The directional order is timed for the two-points-per-lane shapes
    // (the int16 walk's fastest family) and the one-point shapes that walk
    // both directions (the f32 walk's), which gain most from it on large
    // DEMs.
#define TVS_SHAPE(B, U, P, S, D, W)                                                  \
    if (!S) out.push_back(WalkConfig{B, U, true, P, S, D, W});                  \
    if (!S && (W || D)) out.push_back(WalkConfig{B, U, true, P, S, D, W, true}); \
    out.push_back(WalkConfig{B, U, false, P, S, D, W});
    TVS_WALK_SHAPES(TVS_SHAPE)
#undef TVS_SHAPE
    return out;
}

// Block order of a directional launch, per sector slot: the blocks of the
// launch shape (as the walk maps them to points) sorted by the projection of
// their centre on the normal of the sector's band direction.  The band of
// sector s runs along (column, row) = (sin a, cos a), a = s + 0.001 degrees
// (axes.rs:89-100 and the band data it yields: sector 0 steps through rows,
// sector 90 along a row).  Slots past the 180 sectors keep ascending order.
std::vector<uint32_t> directional_walk_order(uint32_t T, uint32_t w, const WalkConfig &cfg,
                                             uint32_t n_slots) {
    const uint32_t bpp = walk_blocks_per_pass(T, cfg);
    const uint32_t B = cfg.block;
    std::vector<double> row(bpp), col(bpp);
    auto at = [&](uint32_t k, double r0, double c0, double r1, double c1) {
        row[k] = 0.5 * (r0 + r1);
        col[k] = 0.5 * (c0 + c1);
    };
    if (cfg.twin) {
        const uint64_t half = T / 2u, n_groups = twin_lane_groups(T, cfg.twin);
        const uint32_t twin_blocks = static_cast<uint32_t>((n_groups + B - 1) / B);
        // Centre of a lane group: one column pair, or (twin 2) the column
        // pairs of rows 2r and 2r + 1.
        auto pair_rc = [&](uint64_t p, double &r, double &cc) {
            const uint64_t rr = p / half;
            if (cfg.twin == 2u) {
                r = 2.0 * static_cast<double>(rr) + 0.5;
                cc = 2.0 * static_cast<double>(p - rr * half) + 1.0;
                return;
            }
            r = static_cast<double>(rr);
            cc = 2.0 * static_cast<double>(p - rr * half) + ((w & rr & 1u) ? 1.5 : 0.5);
        };
        for (uint32_t k = 0; k < bpp; ++k) {
            double r0, c0, r1, c1;
            if (k < twin_blocks) {
                const uint64_t p0 = uint64_t(k) * B;
                pair_rc(p0, r0, c0);
                pair_rc(std::min<uint64_t>(p0 + B, n_groups) - 1u, r1, c1);
            } else {   // each row's left-over column (odd T), one point per lane
                const uint64_t i0 = uint64_t(k - twin_blocks) * B;
                r0 = static_cast<double>(i0);
                r1 = static_cast<double>(std::min<uint64_t>(i0 + B, T) - 1u);
                c0 = c1 = 0.5 * (T - 1);
            }
            at(k, r0, c0, r1, c1);
        }
    } else {
        const uint64_t N = uint64_t(T) * T;
        for (uint32_t k = 0; k < bpp; ++k) {
            const uint64_t t0 = uint64_t(k) * B, t1 = std::min<uint64_t>(t0 + B, N) - 1u;
            at(k, double(t0 / T), double(t0 % T), double(t1 / T), double(t1 % T));
        }
    }
    std::vector<uint32_t> order(static_cast<size_t>(n_slots) * bpp);
    std::vector<double> key(bpp);
    for (uint32_t s = 0; s < n_slots; ++s) {
        uint32_t *o = order.data() + static_cast<size_t>(s) * bpp;
        std::iota(o, o + bpp, 0u);
        if (s >= kSectorSteps) continue;
        const double a = (s + static_cast<double>(kSectorShift)) * (M_PI / 180.0);
        const double ns = std::sin(a), nc = std::cos(a);
        for (uint32_t k = 0; k < bpp; ++k) key[k] = col[k] * nc - row[k] * ns;
        std::stable_sort(o, o + bpp, [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
    }
    return order;
}

}  // namespace tvs

namespace {

tvs_status tfail(tvs_ctx *c, tvs_status st, const std::string &msg) {
    c->err = msg;
    return st;
}

}  // namespace

extern "C" {

tvs_status tvs_set_walk_shape(tvs_ctx *c, const tvs_walk_shape *shape) {
    if (!c || !shape) return TVS_ERR_INVALID_ARGUMENT;
    if (shape->xcd_contiguous > 2u)
        return tfail(c, TVS_ERR_INVALID_ARGUMENT, "xcd_contiguous must be 0, 1 or 2");
    tvs::WalkConfig cfg{shape->block, shape->unroll, shape->xcd_contiguous != 0,
                        shape->pipeline != 0, shape->persistent != 0,
                        shape->both_directions != 0, shape->two_points,
                        shape->xcd_contiguous == 2u};
    if (!tvs::walk_shape_supported(cfg))
        return tfail(c, TVS_ERR_INVALID_ARGUMENT, "walk shape not compiled");
    c->walk_cfg = cfg;
    c->walk_cfg_set = true;
    return TVS_OK;
}

tvs_status tvs_get_walk_shape(const tvs_ctx *c, tvs_walk_shape *out) {
    if (!c || !out) return TVS_ERR_INVALID_ARGUMENT;
    const tvs::WalkConfig cfg = tvs::effective_walk_config(c);
    out->block = cfg.block;
    out->unroll = cfg.unroll;
    out->xcd_contiguous = tvs::xcd_mode(cfg);
    out->pipeline = cfg.pipeline ? 1u : 0u;
    out->persistent = cfg.persistent ? 1u : 0u;
    out->both_directions = cfg.pair ? 1u : 0u;
    out->two_points = cfg.twin;
    out->ms = 0.0f;
    return TVS_OK;
}

size_t tvs_walk_shape_count(void) { return tvs::compiled_walk_shapes().size(); }

tvs_status tvs_autotune(tvs_ctx *c, uint32_t first, uint32_t count, uint32_t reps,
                        tvs_walk_shape *results, size_t max_results, size_t *n_results,
                        tvs_walk_shape *best) {
    tvs::TraceRange trace_range("tvs_autotune");
    if (!c) return TVS_ERR_INVALID_ARGUMENT;
    if (reps == 0) reps = 1;
    const uint32_t N = c->geo.computable_points_count;
    if (N == 0) return tfail(c, TVS_ERR_STATE, "no computable points");
    // Keep the caller's cumulative surfaces.
    float *saved = nullptr;
    double *saved64 = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&saved), N * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&saved64), N * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(saved, c->cum, N * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(saved64, c->cum64, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
        (void)hipFree(saved);
        (void)hipFree(saved64);
        return tfail(c, TVS_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
    }
    const tvs::WalkConfig keep = c->walk_cfg;
    const bool keep_set = c->walk_cfg_set;
    c->walk_cfg_set = true;   // each timed run uses the shape under test
    tvs::WalkConfig winner = keep;
    float winner_ms = 1e30f;
    size_t k = 0;
    tvs_status st = TVS_OK;
    for (const tvs::WalkConfig &cfg : tvs::compiled_walk_shapes()) {
        c->walk_cfg = cfg;
        // Timing runs record no rings (with capture on, every shape would
        // time the one capture kernel) and leave the caller's rings as they
        // are.
        st = tvs::run_sectors(c, first, count, TVS_METHOD_BATCHED, true);   // warm-up
        // The fastest of the timed runs: a run slowed by something else on
        // the device (clocks still rising, another context) does not decide.
        float ms = 1e30f;
        for (uint32_t r = 0; r < reps && st == TVS_OK; ++r) {
            st = tvs::run_sectors(c, first, count, TVS_METHOD_BATCHED, true);
            tvs_timing t;
            if (st == TVS_OK) st = tvs_get_timing(c, &t);
            if (st == TVS_OK) ms = std::min(ms, t.walk_ms);
        }
        if (st != TVS_OK) break;
        if (results && k < max_results)
            results[k] = tvs_walk_shape{cfg.block, cfg.unroll, tvs::xcd_mode(cfg),
                                        cfg.pipeline ? 1u : 0u, cfg.persistent ? 1u : 0u,
                                        cfg.pair ? 1u : 0u, cfg.twin, ms};
        ++k;
        if (ms < winner_ms) {
            winner_ms = ms;
            winner = cfg;
        }
    }
    c->walk_cfg = st == TVS_OK ? winner : keep;
    c->walk_cfg_set = st == TVS_OK ? true : keep_set;
    e = hipMemcpyAsync(c->cum, saved, N * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->cum64, saved64, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(saved);
    (void)hipFree(saved64);
    if (st != TVS_OK) return st;
    if (e != hipSuccess) return tfail(c, TVS_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
    if (n_results) *n_results = k;
    if (best)
        *best = tvs_walk_shape{winner.block, winner.unroll, tvs::xcd_mode(winner),
                               winner.pipeline ? 1u : 0u, winner.persistent ? 1u : 0u,
                               winner.pair ? 1u : 0u, winner.twin, winner_ms};
    return TVS_OK;
}

}  // extern "C"
