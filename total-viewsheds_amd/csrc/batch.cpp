// batch.cpp -- many DEM tiles, all 180 sectors, one device launch.
//
// A batch owns one context per tile (geometry, heights, surfaces).  Band
// data depends only on (width, band size) (band_of_sight.rs:61-173 reads no
// elevation), so tiles of equal (w, B) share the first such tile's step
// tables.  A run is one tile-walk launch over every (tile, sector,
// direction, point block) item, largest bands first, and one reduction
// launch over every point of every tile; the pair is captured once in a
// hipGraph and replayed.
#include <algorithm>
#include <map>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"

struct tvs_batch {
    std::vector<tvs_ctx *> tiles;
    std::vector<uint32_t> owner;     // tile whose band tables a tile uses
    std::vector<uint32_t> order;     // launch order (largest per-item work first)
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {};
    tvs::TileDesc *d_tiles = nullptr;
    uint32_t *d_item_begin = nullptr;
    uint32_t *d_point_begin = nullptr;
    uint32_t n_items = 0, n_points = 0;
    tvs::WalkConfig cfg;
    bool cfg_set = false;        // false: default_walk_config for the tiles' heights
    bool prepared = false;
    bool i16 = false;
    bool ieee_only = false;      // no tile has the fast divides
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    tvs_timing timing{};
    bool ran = false;            // events recorded by a run
    std::string err;
};

namespace {

thread_local std::string g_batch_error;

tvs_status bfail(tvs_batch *b, tvs_status st, const std::string &msg) {
    (b ? b->err : g_batch_error) = msg;
    return st;
}

#define TVS_BHIP(b, expr, what)                                                         \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return bfail((b), _e == hipErrorOutOfMemory ? TVS_ERR_OUT_OF_MEMORY         \
                                                        : TVS_ERR_HIP,                  \
                         std::string(what) + ": " + hipGetErrorString(_e));            \
    } while (0)

tvs_status from_tile(tvs_batch *b, uint32_t i, tvs_status st) {
    if (st != TVS_OK)
        b->err = "tile " + std::to_string(i) + ": " + tvs_last_error(b->tiles[i]);
    return st;
}

// A run is asynchronous on the batch's stream and reads every tile's
// heights, step tables, divide flags and descriptors, and replays the graph
// exec: each call that changes one of them waits for the stream first.
tvs_status quiesce(tvs_batch *b) {
    TVS_BHIP(b, hipStreamSynchronize(b->stream), "wait for the batch's last run");
    return TVS_OK;
}

void drop_graph(tvs_batch *b) {
    // The exec may still be running: destroy it only once the stream is idle.
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->prepared = false;
}

// The shape of an untuned batch: the fastest measured for its height type,
// with its items round-robin over the XCDs (mixed item costs stay
// balanced), or in the directional order when a tile's heights take more
// than kDirectionalDefaultBytes.  Measured per launch, same box,
// round-robin / XCD-contiguous / directional: three 3601 x 3601 int16
// tiles 159.9 / 168.5 / 142.8 ms; sixteen 1201 x 1201 tiles 28.5 / 28.2 /
// 28.7 ms.
tvs::WalkConfig default_batch_config(const tvs_batch *b, bool i16) {
    tvs::WalkConfig cfg = tvs::default_walk_config(i16, false);
    uint64_t largest = 0;
    for (const tvs_ctx *c : b->tiles)
        largest = std::max<uint64_t>(largest,
                                     static_cast<uint64_t>(c->geo.dem_width) * c->geo.dem_width);
    if (largest * (i16 ? 2u : 4u) > tvs::kDirectionalDefaultBytes) {
        cfg.xcd_contiguous = true;
        cfg.directional = true;
    }
    return cfg;
}

// Every tile holds int16 heights (the launch then walks int16; else f32).
bool all_tiles_i16(const tvs_batch *b) {
    bool all = true;
    for (const tvs_ctx *c : b->tiles) all = all && c->has_i16;
    return all;
}

// Upload the descriptors and item/point prefixes for the current shape.
tvs_status prepare(tvs_batch *b) {
    if (b->prepared) return TVS_OK;
    const size_t n = b->tiles.size();
    // One height type per launch: int16 when every tile has it, else f32
    // (every tile holds its heights as f32 too).
    bool all_i16 = true;
    for (tvs_ctx *c : b->tiles) {
        if (!c->has_f32) return bfail(b, TVS_ERR_STATE, "a tile has no elevations");
        all_i16 = all_i16 && c->has_i16;
    }
    b->i16 = all_i16;
    if (!b->cfg_set) b->cfg = default_batch_config(b, b->i16);
    std::vector<tvs::TileDesc> descs(n);
    std::vector<uint32_t> item_begin(n + 1, 0), point_begin(n + 1, 0);
    for (size_t k = 0; k < n; ++k) {
        const uint32_t i = b->order[k];
        tvs_ctx *c = b->tiles[i];
        tvs_ctx *o = b->tiles[b->owner[i]];
        for (uint32_t s = 0; s < tvs::kSectorSteps; ++s)
            if (!c->band_ready[s]) return bfail(b, TVS_ERR_STATE, "tile " + std::to_string(i) + " has no band data");
        tvs::TileDesc d;
        tvs_status st = tvs::prepare_tile(c, 0, tvs::kSectorSteps, 0, b->cfg, &d);
        if (st != TVS_OK) return from_tile(b, i, st);
        if (!b->i16) {
            if (static_cast<uint64_t>(d.w) * d.w > (1ull << 29))
                return bfail(b, TVS_ERR_SIZE, "tile " + std::to_string(i) +
                                                  " too wide for the f32 batched walk");
            d.elev = c->elev_f32;
        }
        // The short divide is checked against the tables the tile walks with,
        // on the tile's stream; the batch walks on its own stream, so the
        // check must be complete before the walk reads its flags.
        st = tvs::ensure_short_div(c, o->tab, 0, tvs::kSectorSteps);
        if (st != TVS_OK) return from_tile(b, i, st);
        TVS_BHIP(b, hipStreamSynchronize(c->stream), "divide check");
        d.off = o->tab.off;
        d.pos_f = o->tab.pos_f;
        d.pos_b = o->tab.pos_b;
        d.dist = o->tab.dist;
        d.rcp = o->tab.rcp;
        d.rcp_lo = o->tab.rcp_lo;
        d.curv = o->tab.curv;
        d.wt = o->tab.wt;
        d.meta = o->tab.meta;
        d.stride = o->tab.stride;
        descs[k] = d;
        item_begin[k + 1] = item_begin[k] + d.blocks_per_pass * 2u * tvs::kSectorSteps;
        point_begin[k + 1] = point_begin[k] + d.N;
    }
    b->n_items = item_begin[n];
    b->n_points = point_begin[n];
    b->ieee_only = std::none_of(descs.begin(), descs.end(),
                                [](const tvs::TileDesc &d) { return d.fast != 0u; });
    (void)hipFree(b->d_tiles);
    (void)hipFree(b->d_item_begin);
    (void)hipFree(b->d_point_begin);
    b->d_tiles = nullptr;
    b->d_item_begin = b->d_point_begin = nullptr;
    TVS_BHIP(b, hipMalloc(reinterpret_cast<void **>(&b->d_tiles), n * sizeof(tvs::TileDesc)), "descriptors");
    TVS_BHIP(b, hipMalloc(reinterpret_cast<void **>(&b->d_item_begin), (n + 1) * sizeof(uint32_t)), "descriptors");
    TVS_BHIP(b, hipMalloc(reinterpret_cast<void **>(&b->d_point_begin), (n + 1) * sizeof(uint32_t)), "descriptors");
    TVS_BHIP(b, hipMemcpy(b->d_tiles, descs.data(), n * sizeof(tvs::TileDesc), hipMemcpyHostToDevice), "descriptors");
    TVS_BHIP(b, hipMemcpy(b->d_item_begin, item_begin.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice), "descriptors");
    TVS_BHIP(b, hipMemcpy(b->d_point_begin, point_begin.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice), "descriptors");
    b->prepared = true;
    return TVS_OK;
}

tvs_status enqueue(tvs_batch *b) {
    TVS_BHIP(b, tvs::launch_walk_tiles(b->stream, b->d_tiles, b->d_item_begin,
                                       static_cast<uint32_t>(b->tiles.size()), b->n_items,
                                       b->i16, false, b->ieee_only, b->cfg),
             "tile walk");
    TVS_BHIP(b, tvs::launch_reduce_tiles(b->stream, b->d_tiles, b->d_point_begin,
                                         static_cast<uint32_t>(b->tiles.size()), b->n_points),
             "tile reduction");
    return TVS_OK;
}

}  // namespace

extern "C" {

const char *tvs_batch_last_error(const tvs_batch *b) {
    return b ? b->err.c_str() : g_batch_error.c_str();
}

tvs_status tvs_batch_create(const tvs_params *tiles, uint32_t n_tiles, tvs_batch **out) {
    if (!tiles || !out || n_tiles == 0)
        return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "no tiles");
    *out = nullptr;
    tvs_batch *b = new (std::nothrow) tvs_batch();
    if (!b) return bfail(nullptr, TVS_ERR_OUT_OF_MEMORY, "host allocation");
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> first_of;
    for (uint32_t i = 0; i < n_tiles; ++i) {
        tvs_params p = tiles[i];
        p.ring_capture_sectors = 0;
        tvs_ctx *c = nullptr;
        const tvs_status st = tvs_create(&p, &c);
        if (st != TVS_OK) {
            const std::string msg = "tile " + std::to_string(i) + ": " + tvs_last_error(nullptr);
            tvs_batch_destroy(b);
            return bfail(nullptr, st, msg);
        }
        b->tiles.push_back(c);
        const auto key = std::make_pair(c->geo.dem_width, c->geo.band_size);
        auto it = first_of.find(key);
        b->owner.push_back(it == first_of.end() ? i : it->second);
        if (it == first_of.end()) first_of[key] = i;
    }
    // Largest per-item work (steps per band) first, so the long items start
    // early and short ones fill the end of the launch.
    b->order.resize(n_tiles);
    std::iota(b->order.begin(), b->order.end(), 0u);
    std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) {
        return b->tiles[x]->geo.band_deltas_size > b->tiles[y]->geo.band_deltas_size;
    });
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipEventCreate(&b->ev[k]);
    if (e != hipSuccess) {
        tvs_batch_destroy(b);
        return bfail(nullptr, TVS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    }
    *out = b;
    return TVS_OK;
}

tvs_status tvs_batch_destroy(tvs_batch *b) {
    if (!b) return TVS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (tvs_ctx *c : b->tiles) tvs_destroy(c);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_item_begin) (void)hipFree(b->d_item_begin);
    if (b->d_point_begin) (void)hipFree(b->d_point_begin);
    for (auto &ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TVS_OK;
}

tvs_status tvs_batch_tile_geometry(const tvs_batch *b, uint32_t tile, tvs_geometry *out) {
    if (!b || !out || tile >= b->tiles.size())
        return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "tile index");
    *out = b->tiles[tile]->geo;
    return TVS_OK;
}

tvs_status tvs_batch_upload_i16(tvs_batch *b, uint32_t tile, const int16_t *elevations,
                                size_t count) {
    if (!b || tile >= b->tiles.size()) return bfail(b, TVS_ERR_INVALID_ARGUMENT, "tile index");
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    drop_graph(b);
    return from_tile(b, tile, tvs_upload_dem_i16(b->tiles[tile], elevations, count));
}

tvs_status tvs_batch_upload_f32(tvs_batch *b, uint32_t tile, const float *elevations,
                                size_t count) {
    if (!b || tile >= b->tiles.size()) return bfail(b, TVS_ERR_INVALID_ARGUMENT, "tile index");
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    drop_graph(b);
    return from_tile(b, tile, tvs_upload_dem_f32(b->tiles[tile], elevations, count));
}

tvs_status tvs_batch_compute_bands(tvs_batch *b) {
    tvs::TraceRange trace_range("tvs_batch_compute_bands");
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    float ms = 0.0f;
    for (uint32_t i = 0; i < b->tiles.size(); ++i) {
        if (b->owner[i] != i) continue;
        const tvs_status st = tvs_compute_bands(b->tiles[i], 0, tvs::kSectorSteps);
        if (st != TVS_OK) return from_tile(b, i, st);
        ms += b->tiles[i]->timing.bands_ms;
    }
    for (uint32_t i = 0; i < b->tiles.size(); ++i) {
        const uint32_t o = b->owner[i];
        tvs_ctx *c = b->tiles[i];
        std::fill(c->short_div_checked.begin(), c->short_div_checked.end(), 0);   // new tables
        if (o == i) continue;
        c->meta_host = b->tiles[o]->meta_host;
        const tvs_status st = tvs::validate_slots(c, 0, tvs::kSectorSteps);
        if (st != TVS_OK) return from_tile(b, i, st);
    }
    drop_graph(b);
    b->timing.bands_ms = ms;
    return TVS_OK;
}

tvs_status tvs_batch_set_walk_shape(tvs_batch *b, const tvs_walk_shape *shape) {
    if (!b || !shape) return bfail(b, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (shape->xcd_contiguous > 2u)
        return bfail(b, TVS_ERR_INVALID_ARGUMENT, "xcd_contiguous must be 0, 1 or 2");
    tvs::WalkConfig cfg{shape->block, shape->unroll, shape->xcd_contiguous != 0,
                        shape->pipeline != 0, shape->persistent != 0,
                        shape->both_directions != 0, shape->two_points,
                        shape->xcd_contiguous == 2u};
    if (!tvs::walk_shape_supported(cfg))
        return bfail(b, TVS_ERR_INVALID_ARGUMENT, "walk shape not compiled");
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    b->cfg = cfg;
    b->cfg_set = true;
    drop_graph(b);
    return TVS_OK;
}

tvs_status tvs_batch_get_walk_shape(const tvs_batch *b, tvs_walk_shape *out) {
    if (!b || !out) return TVS_ERR_INVALID_ARGUMENT;
    const tvs::WalkConfig cfg = b->cfg_set ? b->cfg : default_batch_config(b, all_tiles_i16(b));
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

tvs_status tvs_batch_autotune(tvs_batch *b, uint32_t reps, tvs_walk_shape *results,
                              size_t max_results, size_t *n_results, tvs_walk_shape *best) {
    tvs::TraceRange trace_range("tvs_batch_autotune");
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    if (reps == 0) reps = 1;
    // Only the walk launch is timed: it writes the per-sector band surfaces
    // (scratch), never the cumulative surfaces, so the caller's totals stay.
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    const tvs::WalkConfig keep = b->cfg;
    const bool keep_set = b->cfg_set;
    b->cfg_set = true;   // prepare() must use the shape under test
    tvs::WalkConfig winner = keep;
    float winner_ms = 1e30f;
    size_t k = 0;
    tvs_status st = TVS_OK;
    for (const tvs::WalkConfig &cfg : tvs::compiled_walk_shapes()) {
        b->cfg = cfg;
        drop_graph(b);
        st = prepare(b);
        if (st != TVS_OK) break;
        auto launch = [&]() {
            return tvs::launch_walk_tiles(b->stream, b->d_tiles, b->d_item_begin,
                                          static_cast<uint32_t>(b->tiles.size()), b->n_items,
                                          b->i16, false, b->ieee_only, b->cfg);
        };
        hipError_t e = launch();   // warm-up
        // The fastest of the timed launches (as tvs_autotune).
        float ms = 1e30f;
        for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
            e = hipEventRecord(b->ev[0], b->stream);
            if (e == hipSuccess) e = launch();
            if (e == hipSuccess) e = hipEventRecord(b->ev[1], b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(b->ev[1]);
            float one = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&one, b->ev[0], b->ev[1]);
            if (e == hipSuccess) ms = std::min(ms, one);
        }
        if (e != hipSuccess) {
            st = bfail(b, TVS_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
            break;
        }
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
    b->cfg = st == TVS_OK ? winner : keep;
    b->cfg_set = st == TVS_OK ? true : keep_set;
    drop_graph(b);
    b->ran = false;   // the run events now hold tuner timings
    if (st != TVS_OK) return st;
    if (n_results) *n_results = k;
    if (best)
        *best = tvs_walk_shape{winner.block, winner.unroll, tvs::xcd_mode(winner),
                               winner.pipeline ? 1u : 0u, winner.persistent ? 1u : 0u,
                               winner.pair ? 1u : 0u, winner.twin, winner_ms};
    return TVS_OK;
}

tvs_status tvs_batch_set_short_divide(tvs_batch *b, uint32_t enable) {
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    for (uint32_t i = 0; i < b->tiles.size(); ++i) {
        const tvs_status st = tvs_set_short_divide(b->tiles[i], enable);
        if (st != TVS_OK) return from_tile(b, i, st);
    }
    drop_graph(b);
    return TVS_OK;
}

tvs_status tvs_batch_set_half_heights(tvs_batch *b, uint32_t enable) {
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    const tvs_status q = quiesce(b);
    if (q != TVS_OK) return q;
    for (uint32_t i = 0; i < b->tiles.size(); ++i) {
        const tvs_status st = tvs_set_half_heights(b->tiles[i], enable);
        if (st != TVS_OK) return from_tile(b, i, st);
    }
    drop_graph(b);
    return TVS_OK;
}

tvs_status tvs_batch_run(tvs_batch *b, uint32_t use_graph) {
    tvs::TraceRange trace_range("tvs_batch_run");
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    tvs_status st = prepare(b);
    if (st != TVS_OK) return st;
    uint64_t steps = 0;
    for (tvs_ctx *c : b->tiles)
        steps += 180ull * 2ull * c->geo.computable_points_count * c->geo.band_deltas_size;
    b->timing = tvs_timing{0.0f, 0.0f, b->timing.bands_ms, 1u, steps};
    b->ran = true;
    if (use_graph) {
        if (!b->exec) {
            TVS_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "capture");
            const tvs_status es = enqueue(b);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (es != TVS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return es;
            }
            TVS_BHIP(b, ce, "capture");
            b->graph = g;
            TVS_BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph");
        }
        TVS_BHIP(b, hipEventRecord(b->ev[0], b->stream), "event");
        TVS_BHIP(b, hipGraphLaunch(b->exec, b->stream), "graph launch");
        TVS_BHIP(b, hipEventRecord(b->ev[1], b->stream), "event");
        TVS_BHIP(b, hipEventRecord(b->ev[2], b->stream), "event");
        return TVS_OK;
    }
    TVS_BHIP(b, hipEventRecord(b->ev[0], b->stream), "event");
    TVS_BHIP(b, tvs::launch_walk_tiles(b->stream, b->d_tiles, b->d_item_begin,
                                       static_cast<uint32_t>(b->tiles.size()), b->n_items,
                                       b->i16, false, b->ieee_only, b->cfg),
             "tile walk");
    TVS_BHIP(b, hipEventRecord(b->ev[1], b->stream), "event");
    TVS_BHIP(b, tvs::launch_reduce_tiles(b->stream, b->d_tiles, b->d_point_begin,
                                         static_cast<uint32_t>(b->tiles.size()), b->n_points),
             "tile reduction");
    TVS_BHIP(b, hipEventRecord(b->ev[2], b->stream), "event");
    return TVS_OK;
}

tvs_status tvs_batch_sync(tvs_batch *b) {
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    TVS_BHIP(b, hipStreamSynchronize(b->stream), "sync");
    return TVS_OK;
}

tvs_status tvs_batch_reset_surfaces(tvs_batch *b) {
    if (!b) return bfail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null batch");
    TVS_BHIP(b, hipStreamSynchronize(b->stream), "sync");
    for (uint32_t i = 0; i < b->tiles.size(); ++i) {
        tvs_status st = tvs_reset_surfaces(b->tiles[i]);
        if (st == TVS_OK) st = tvs_sync(b->tiles[i]);
        if (st != TVS_OK) return from_tile(b, i, st);
    }
    return TVS_OK;
}

tvs_status tvs_batch_read_surfaces(tvs_batch *b, uint32_t tile, float *out, size_t n) {
    if (!b || tile >= b->tiles.size()) return bfail(b, TVS_ERR_INVALID_ARGUMENT, "tile index");
    TVS_BHIP(b, hipStreamSynchronize(b->stream), "sync");
    return from_tile(b, tile, tvs_read_surfaces(b->tiles[tile], out, n));
}

tvs_status tvs_batch_read_surfaces_f64(tvs_batch *b, uint32_t tile, double *out, size_t n) {
    if (!b || tile >= b->tiles.size()) return bfail(b, TVS_ERR_INVALID_ARGUMENT, "tile index");
    TVS_BHIP(b, hipStreamSynchronize(b->stream), "sync");
    return from_tile(b, tile, tvs_read_surfaces_f64(b->tiles[tile], out, n));
}

tvs_status tvs_batch_get_timing(tvs_batch *b, tvs_timing *out) {
    if (!b || !out) return bfail(b, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (!b->ran) {
        *out = b->timing;
        return TVS_OK;
    }
    TVS_BHIP(b, hipEventSynchronize(b->ev[2]), "event");
    float ms = 0.0f;
    TVS_BHIP(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]), "event time");
    b->timing.walk_ms = ms;
    TVS_BHIP(b, hipEventElapsedTime(&ms, b->ev[1], b->ev[2]), "event time");
    b->timing.reduce_ms = ms;
    *out = b->timing;
    return TVS_OK;
}

}  // extern "C"
