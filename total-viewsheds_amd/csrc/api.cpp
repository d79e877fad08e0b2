// api.cpp -- the extern "C" boundary (include/tvs.h): context lifetime,
// uploads, band data, walk dispatch, readback, errors.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"

namespace tvs {
uint32_t max_band_samples();
}

namespace {

thread_local std::string g_error;  // errors with no context to hold them

tvs_status fail(tvs_ctx *c, tvs_status st, const std::string &msg) {
    (c ? c->err : g_error) = msg;
    return st;
}

tvs_status hip_fail(tvs_ctx *c, hipError_t e, const char *what) {
    const tvs_status st =
        (e == hipErrorOutOfMemory) ? TVS_ERR_OUT_OF_MEMORY : TVS_ERR_HIP;
    return fail(c, st, std::string(what) + ": " + hipGetErrorString(e));
}

#define TVS_HIP(c, expr, what)                                  \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return hip_fail((c), _e, (what)); \
    } while (0)

template <typename T>
hipError_t dev_alloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <typename T>
void dev_free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Bounds on |dh| = |fl32(h - fl32(hp + obs))| over every pair of integer
// heights h, hp in [lo, hi] (a superset of the DEM's values).  The fast
// divide needs every non-zero |dh| in [2^-40, 2^40].
bool fast_divide_holds(int64_t lo, int64_t hi, float obs, float *min_nz,
                       float *max_abs) {
    if (hi < lo || hi - lo > (int64_t(1) << 20)) return false;
    float mn = INFINITY, mx = 0.0f;
    for (int64_t hp = lo; hp <= hi; ++hp) {
        const float h0 = static_cast<float>(hp) + obs;
        if (!std::isfinite(h0)) return false;
        mx = std::max(mx, std::fabs(static_cast<float>(hi) - h0));
        mx = std::max(mx, std::fabs(static_cast<float>(lo) - h0));
        const int64_t f = static_cast<int64_t>(std::floor(static_cast<double>(h0)));
        const int64_t cand[6] = {f - 1, f, f + 1, f + 2, lo, hi};
        for (int64_t h : cand) {
            if (h < lo || h > hi) continue;
            const float dh = static_cast<float>(h) - h0;
            if (dh != 0.0f) mn = std::min(mn, std::fabs(dh));
        }
    }
    *min_nz = mn;
    *max_abs = mx;
    return mn >= 0x1p-40f && mx <= 0x1p40f;
}

tvs_status check_band_geometry(tvs_ctx *c) {
    const tvs_geometry &g = c->geo;
    const uint32_t middle = g.band_samples / 2u;
    const uint32_t pov_id = g.dem_size / 2u;
    // band_of_sight.rs:103 (usize underflow), 113 (window past the DEM),
    // 162 (band_id + 1 past the samples), 165 (i past band_size - 1).
    if (middle > pov_id || pov_id - middle + g.band_samples > g.dem_size ||
        2u * g.band_size >= g.band_samples || g.band_size < 2u ||
        2u * g.band_size - middle > g.band_deltas_size)
        return fail(c, TVS_ERR_BAND_TOO_SHORT,
                    "band template does not fit (band_size " + std::to_string(g.band_size) +
                        ", band_samples " + std::to_string(g.band_samples) +
                        "): the reference panics here (band_of_sight.rs:103-168)");
    if (g.band_samples > tvs::max_band_samples())
        return fail(c, TVS_ERR_INVALID_ARGUMENT,
                    "band_samples " + std::to_string(g.band_samples) +
                        " exceeds the device band compile limit " +
                        std::to_string(tvs::max_band_samples()));
    return TVS_OK;
}

// kernel.rs:131, 157-164: every band point of every observer must index the
// elevation array.
tvs_status check_bounds(tvs_ctx *c, uint32_t slot) {
    const tvs_geometry &g = c->geo;
    const tvs::SectorMeta &m = c->meta_host[slot];
    if (g.computable_points_count == 0) return TVS_OK;
    const int64_t N = g.computable_points_count, T = g.tvs_width, L = g.max_los_as_points,
                  w = g.dem_width;
    const int64_t pov_min = L * w + L;
    const int64_t pov_max = ((N - 1) / T + L) * w + ((N - 1) % T + L);
    const int64_t lo = std::min<int64_t>(m.min_off, -static_cast<int64_t>(m.max_off));
    const int64_t hi = std::max<int64_t>(m.max_off, -static_cast<int64_t>(m.min_off));
    if (pov_min + lo < 0 || pov_max + hi > w * w - 1)
        return fail(c, TVS_ERR_BAND_OUT_OF_DEM,
                    "sector slot " + std::to_string(slot) +
                        ": band leaves the elevation array (kernel.rs:164 would panic)");
    return TVS_OK;
}

tvs_status ensure_sort_scratch(tvs_ctx *c) {
    tvs::SortScratch &s = c->sort;
    const size_t n = static_cast<size_t>(c->geo.dem_size);
    if (s.n == n && s.key_a) return TVS_OK;
    const uint32_t tiles = static_cast<uint32_t>((n + 4095) / 4096);
    TVS_HIP(c, dev_alloc(&s.key_a, n), "sort keys");
    TVS_HIP(c, dev_alloc(&s.key_b, n), "sort keys");
    TVS_HIP(c, dev_alloc(&s.id_a, n), "sort ids");
    TVS_HIP(c, dev_alloc(&s.id_b, n), "sort ids");
    TVS_HIP(c, dev_alloc(&s.sight_map, n), "sight map");
    TVS_HIP(c, dev_alloc(&s.dist_f32, n), "axes distances");
    TVS_HIP(c, dev_alloc(&s.hist, static_cast<size_t>(256) * (tiles + 1)), "sort histogram");
    s.n = n;
    s.tiles = tiles;
    return TVS_OK;
}

void free_sort_scratch(tvs::SortScratch &s) {
    dev_free(s.key_a);
    dev_free(s.key_b);
    dev_free(s.id_a);
    dev_free(s.id_b);
    dev_free(s.sight_map);
    dev_free(s.dist_f32);
    dev_free(s.hist);
    s.n = 0;
}

tvs_status read_meta(tvs_ctx *c, uint32_t first, uint32_t count) {
    TVS_HIP(c, hipStreamSynchronize(c->stream), "band tables");
    TVS_HIP(c, hipMemcpy(c->meta_host.data() + first, c->tab.meta + first,
                         count * sizeof(tvs::SectorMeta), hipMemcpyDeviceToHost),
            "band table metadata");
    return tvs::validate_slots(c, first, count);
}

tvs_status ensure_ring(tvs_ctx *c, size_t slots) {
    if (c->ring_slots >= slots && c->ring) return TVS_OK;
    dev_free(c->ring);
    c->ring_slots = 0;
    const size_t per = 2ull * c->geo.computable_points_count * c->rings_per_band;
    TVS_HIP(c, dev_alloc(&c->ring, per * slots), "ring buffer");
    c->ring_slots = slots;
    return TVS_OK;
}

// Every height difference fl32(h - fl32(hp + obs)) for integer h, hp in
// [lo, hi] (a superset of what the walk can meet).  h0 = fl32(hp + obs)
// equals hp + o with o taking a few values (one per binade of hp + obs), and
// h - h0 = (h - hp) - o exactly, so the set is {fl32(D - o)} over the
// integer differences D in [lo - hi, hi - lo] and the distinct o; each
// value is computed the way the walk computes it, from one hp per o.
std::vector<float> height_differences(int64_t lo, int64_t hi, float obs) {
    std::vector<int64_t> reps;
    std::vector<double> offsets;
    for (int64_t hp = lo; hp <= hi; ++hp) {
        const float h0 = static_cast<float>(hp) + obs;
        const double o = static_cast<double>(h0) - static_cast<double>(hp);
        if (std::find(offsets.begin(), offsets.end(), o) == offsets.end()) {
            offsets.push_back(o);
            reps.push_back(hp);
        }
    }
    std::vector<float> out;
    out.reserve(reps.size() * static_cast<size_t>(2 * (hi - lo) + 1));
    for (int64_t hp : reps) {
        const float h0 = static_cast<float>(hp) + obs;
        for (int64_t dd = lo - hi; dd <= hi - lo; ++dd)
            out.push_back(static_cast<float>(hp + dd) - h0);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

// Largest domain the short-divide check takes (values); beyond it the
// split-reciprocal divide is used everywhere.
constexpr size_t kMaxDivDomain = size_t(1) << 22;

// Any f32 heights: the four-operation divide's proof needs every
// non-zero |dh| in [2^-40, 2^40], dh = fl32(h - h0) with h a DEM value and
// h0 = fl32(hp + obs).  If every h and every h0 is 0 or has a magnitude in
// [2^-17, 2^38], both are multiples of 2^-40 (an f32 of magnitude >= 2^-17
// has an ulp >= 2^-40), so a non-zero difference, and its rounding, is at
// least 2^-40 in magnitude, and |dh| <= 2^39.  NaN and infinities fail.
bool spaced_height(float x) {
    const float a = std::fabs(x);
    return x == 0.0f || (a >= 0x1p-17f && a <= 0x1p38f);
}

bool spaced_divide_holds(const float *elev, size_t count, float obs) {
    for (size_t i = 0; i < count; ++i)
        if (!spaced_height(elev[i]) || !spaced_height(elev[i] + obs)) return false;
    return true;
}

// `integral`: heights are integers in [lo, hi] (the short divide's domain
// can be enumerated).  `spaced_ok`: every height and observer height passes
// spaced_height, which also admits the four-operation divide (integer
// heights whose range is too wide for fast_divide_holds' scan among them).
tvs_status upload_common(tvs_ctx *c, bool integral, bool spaced_ok, int64_t lo,
                         int64_t hi) {
    ++c->state_gen;
    float mn = 0.0f, mx = 0.0f;
    c->fast_div_ok =
        integral && fast_divide_holds(lo, hi, c->params.observer_height, &mn, &mx);
    if (!c->fast_div_ok && spaced_ok) {
        c->fast_div_ok = true;
        mn = 0x1p-40f;   // bounds, not the extremes
        mx = 0x1p39f;
    }
    c->min_abs_dh = mn;
    c->max_abs_dh = mx;
    std::fill(c->short_div_checked.begin(), c->short_div_checked.end(), 0);
    dev_free(c->div_domain);
    c->div_domain_n = 0;
    // The short divide is checked over every height difference the DEM can
    // produce, enumerable only for integer heights.
    if (!c->fast_div_ok || !integral) return TVS_OK;
    const std::vector<float> dom = height_differences(lo, hi, c->params.observer_height);
    if (dom.empty() || dom.size() > kMaxDivDomain) return TVS_OK;
    TVS_HIP(c, dev_alloc(&c->div_domain, dom.size()), "height differences");
    TVS_HIP(c, hipMemcpy(c->div_domain, dom.data(), dom.size() * sizeof(float),
                         hipMemcpyHostToDevice),
            "height differences");
    c->div_domain_n = static_cast<uint32_t>(dom.size());
    return TVS_OK;
}

}  // namespace

namespace tvs {

// Bounds check of slots [first, first+count) against this context's
// geometry (meta_host already filled); marks them ready.
tvs_status validate_slots(tvs_ctx *c, uint32_t first, uint32_t count) {
    ++c->state_gen;
    for (uint32_t s = first; s < first + count; ++s) {
        const tvs_status st = check_bounds(c, s);
        if (st != TVS_OK) {
            c->band_ready[s] = 0;
            return st;
        }
        c->band_ready[s] = 1;
    }
    return TVS_OK;
}

WalkConfig effective_walk_config(const tvs_ctx *c) {
    if (c->walk_cfg_set) return c->walk_cfg;
    WalkConfig cfg = default_walk_config(c->has_i16, true);
    cfg.directional = static_cast<uint64_t>(c->geo.dem_width) * c->geo.dem_width *
                          (c->has_i16 ? 2u : 4u) > kDirectionalDefaultBytes;
    return cfg;
}

// Verifies the short divide for the unchecked slots of [first, first +
// count) of context c against step tables `tab` (c's own, or the band
// owner's in a batch).  No-op when the short divide is off or unavailable.
tvs_status ensure_short_div(tvs_ctx *c, const StepTables &tab, uint32_t first, uint32_t count) {
    if (!c->short_div || !c->fast_div_ok || !c->div_domain) return TVS_OK;
    uint32_t s = first;
    while (s < first + count) {
        if (c->short_div_checked[s]) {
            ++s;
            continue;
        }
        uint32_t e = s;
        while (e < first + count && !c->short_div_checked[e]) ++e;
        TVS_HIP(c, launch_fill_u32(c->stream, c->short_div_bad + s, e - s, 0u), "divide check");
        TVS_HIP(c, launch_short_div_check(c->stream, tab, s, e - s, c->div_domain,
                                     c->div_domain_n, c->short_div_bad),
                "divide check");
        for (uint32_t k = s; k < e; ++k) c->short_div_checked[k] = 1;
        s = e;
    }
    return TVS_OK;
}

// Descriptor of this context's DEM for a batched launch over sector slots
// [first, first + count); grows the band-surface buffer as needed.  With
// `allocate` false nothing is allocated: the caller sets `partial` itself
// and the shape must not be directional (its block order is device memory
// built on demand).
tvs_status prepare_tile(tvs_ctx *c, uint32_t first, uint32_t count, uint32_t capture,
                        const WalkConfig &cfg, TileDesc *out, bool allocate) {
    const uint32_t N = c->geo.computable_points_count;
    if (!walk_shape_supported(cfg))
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "walk shape not compiled");
    // Heights are addressed with 32-bit byte offsets from a resource that
    // starts up to w*w - 1 elements below the DEM (step positions are at
    // most 4 (w*w - 1) int16 bytes, twice that for f32 heights; DEMs with a
    // second copy are far below these limits).
    const uint64_t w2 = static_cast<uint64_t>(c->geo.dem_width) * c->geo.dem_width;
    if (w2 > (c->has_i16 ? (1ull << 30) : (1ull << 29)))
        return fail(c, TVS_ERR_SIZE,
                    "DEM too wide for the batched walk's 32-bit height offsets (int16 up to "
                    "32768 points wide, f32 up to 23170); use TVS_METHOD_PLAIN");
    if (!allocate && cfg.directional)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "directional block order needs allocation");
    if (allocate && c->partial_cap < count) {
        dev_free(c->partial);
        c->partial_cap = 0;
        TVS_HIP(c, dev_alloc(&c->partial, static_cast<size_t>(count) * 2ull * N),
                "band surfaces");
        c->partial_cap = count;
    }
    TileDesc d{};
    d.elev = c->has_i16 ? static_cast<const void *>(c->elev_i16)
                        : static_cast<const void *>(c->elev_f32);
    d.span = static_cast<uint32_t>(dem_span(w2));
    d.copy_offset = static_cast<uint32_t>(dem_copy_offset(w2));
    d.pov_shift = pair_parity(c->geo.dem_width, c->geo.max_los_as_points);
    d.elev_h16 = c->has_i16 && c->has_h16 && c->half_heights ? c->elev_h16 : nullptr;
    d.h16_offset = c->h16_offset;
    d.off = c->tab.off;
    d.pos_f = c->tab.pos_f;
    d.pos_b = c->tab.pos_b;
    d.dist = c->tab.dist;
    d.rcp = c->tab.rcp;
    d.rcp_lo = c->tab.rcp_lo;
    d.curv = c->tab.curv;
    d.wt = c->tab.wt;
    d.meta = c->tab.meta;
    d.partial = c->partial;
    d.cum = c->cum;
    d.cum64 = c->cum64;
    d.ring = capture ? c->ring : nullptr;
    d.overflow = c->overflow;
    d.stride = c->tab.stride;
    d.n_steps = c->geo.band_deltas_size;
    d.first_slot = first;
    d.count = count;
    d.N = N;
    d.T = c->geo.tvs_width;
    d.L = c->geo.max_los_as_points;
    d.w = c->geo.dem_width;
    d.obs = c->params.observer_height;
    d.blocks_per_pass = walk_blocks_per_pass(c->geo.tvs_width, cfg);
    d.R = c->rings_per_band;
    d.capture_slots = capture;
    d.fast = c->fast_div_ok ? 1u : 0u;
    d.short_div_bad = (c->short_div && c->fast_div_ok && c->div_domain) ? c->short_div_bad : nullptr;
    d.ring_mode = kRingRecords;
    d.ring_counts = nullptr;
    d.ring_offsets = nullptr;
    d.ring_events = nullptr;
    d.order = nullptr;
    if (allocate && cfg.directional && cfg.xcd_contiguous && !cfg.persistent) {
        const uint64_t key = (1ull << 63) | (uint64_t(cfg.twin) << 48) |
                             (uint64_t(cfg.block) << 32) | (uint64_t(d.w) << 16) | d.T;
        if (c->order_key != key) {
            const std::vector<uint32_t> order =
                directional_walk_order(d.T, d.w, cfg, kBandSlots);
            dev_free(c->d_order);
            c->order_key = 0;
            TVS_HIP(c, dev_alloc(&c->d_order, order.size()), "walk order");
            TVS_HIP(c, hipMemcpy(c->d_order, order.data(), order.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice),
                    "walk order");
            c->order_key = key;
        }
        d.order = c->d_order;
    }
    *out = d;
    return TVS_OK;
}

}  // namespace tvs

extern "C" {

const char *tvs_status_string(tvs_status st) {
    switch (st) {
        case TVS_OK: return "ok";
        case TVS_ERR_INVALID_ARGUMENT: return "invalid argument";
        case TVS_ERR_LOS_TOO_LONG: return "max line of sight too long";
        case TVS_ERR_BAND_TOO_SHORT: return "band template does not fit";
        case TVS_ERR_BAND_OUT_OF_DEM: return "band leaves the DEM";
        case TVS_ERR_SIZE: return "size mismatch";
        case TVS_ERR_STATE: return "call out of order";
        case TVS_ERR_NO_DEVICE: return "no HIP device";
        case TVS_ERR_OUT_OF_MEMORY: return "out of device memory";
        case TVS_ERR_HIP: return "HIP runtime error";
        case TVS_ERR_IO: return "I/O error";
    }
    return "unknown status";
}

const char *tvs_last_error(const tvs_ctx *c) {
    return c ? c->err.c_str() : g_error.c_str();
}

tvs_status tvs_create(const tvs_params *p, tvs_ctx **out) {
    if (!p || !out) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tvs_params params = *p;
    if (params.max_line_of_sight == 0)
        params.max_line_of_sight =
            tvs_default_max_line_of_sight(params.dem_width, static_cast<double>(params.scale));
    tvs_geometry g;
    tvs_status st = tvs::dem_geometry(params.dem_width, params.scale,
                                      params.max_line_of_sight, &g);
    if (st == TVS_ERR_LOS_TOO_LONG)
        return fail(nullptr, st,
                    "The maximum line of sight (" + std::to_string(params.max_line_of_sight) +
                        "m) is longer than the maximum distance that any point can "
                        "completely see (" +
                        std::to_string(static_cast<double>(params.dem_width / 2u) *
                                       static_cast<double>(params.scale)) +
                        "m).");
    if (st != TVS_OK) return fail(nullptr, st, "invalid DEM width or scale");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, TVS_ERR_NO_DEVICE, "no HIP device visible");
    tvs_ctx *c = new (std::nothrow) tvs_ctx();
    if (!c) return fail(nullptr, TVS_ERR_OUT_OF_MEMORY, "host allocation");
    c->params = params;
    c->geo = g;
    c->rings_per_band = tvs_ring_count_per_band(params.rings_per_km, params.max_line_of_sight);
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; i < 6 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    const uint32_t N = g.computable_points_count;
    if (e == hipSuccess) e = dev_alloc(&c->cum, N);
    if (e == hipSuccess) e = dev_alloc(&c->cum64, N);
    if (e == hipSuccess) e = dev_alloc(&c->overflow, 1);
    if (e == hipSuccess) e = dev_alloc(&c->d_tile, 1);
    if (e == hipSuccess) e = dev_alloc(&c->d_axes, 4u * tvs::kSectorSteps);
    if (e == hipSuccess) e = dev_alloc(&c->d_select_status, tvs::kSectorSteps);
    if (e == hipSuccess) {
        std::vector<double> axes(4u * tvs::kSectorSteps);
        for (uint32_t s = 0; s < tvs::kSectorSteps; ++s) {
            tvs::sector_sin_cos(s, false, &axes[4 * s], &axes[4 * s + 1]);
            tvs::sector_sin_cos(s, true, &axes[4 * s + 2], &axes[4 * s + 3]);
        }
        e = hipMemcpy(c->d_axes, axes.data(), axes.size() * sizeof(double),
                      hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = dev_alloc(&c->d_begin, 4);
    c->tab.stride = std::max<uint32_t>(64u, (g.band_deltas_size + 63u) / 64u * 64u);
    const size_t tab_n = static_cast<size_t>(tvs::kBandSlots) * c->tab.stride;
    if (e == hipSuccess) e = dev_alloc(&c->tab.delta, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.dist, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.off, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.pos_f, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.pos_b, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.rcp, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.rcp_lo, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.curv, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.wt, tab_n);
    if (e == hipSuccess) e = dev_alloc(&c->tab.meta, tvs::kBandSlots);
    if (e == hipSuccess) e = dev_alloc(&c->short_div_bad, tvs::kBandSlots);
    if (e == hipSuccess) e = hipMemsetAsync(c->cum, 0, sizeof(float) * std::max(N, 1u), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->cum64, 0, sizeof(double) * std::max(N, 1u), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->overflow, 0, sizeof(unsigned long long), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        const tvs_status s2 = hip_fail(nullptr, e, "tvs_create");
        tvs_destroy(c);
        return s2;
    }
    c->band_ready.assign(tvs::kBandSlots, 0);
    c->short_div_checked.assign(tvs::kBandSlots, 0);
    c->meta_host.assign(tvs::kBandSlots, tvs::SectorMeta{});
    *out = c;
    return TVS_OK;
}

tvs_status tvs_destroy(tvs_ctx *c) {
    if (!c) return TVS_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dev_free(c->elev_f32);
    dev_free(c->elev_i16);
    dev_free(c->elev_h16);
    dev_free(c->tab.delta);
    dev_free(c->tab.dist);
    dev_free(c->tab.off);
    dev_free(c->tab.pos_f);
    dev_free(c->tab.pos_b);
    dev_free(c->tab.rcp);
    dev_free(c->tab.rcp_lo);
    dev_free(c->tab.curv);
    dev_free(c->tab.wt);
    dev_free(c->tab.meta);
    dev_free(c->short_div_bad);
    dev_free(c->div_domain);
    dev_free(c->cum);
    dev_free(c->cum64);
    dev_free(c->partial);
    dev_free(c->ring);
    dev_free(c->overflow);
    dev_free(c->d_tile);
    dev_free(c->d_ring_counts);
    dev_free(c->d_ring_offsets);
    dev_free(c->d_ring_events);
    dev_free(c->d_order);
    dev_free(c->d_obs);
    dev_free(c->d_axes);
    dev_free(c->d_select_status);
    dev_free(c->d_begin);
    free_sort_scratch(c->sort);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t h : c->walk_hist)
        if (h) (void)hipEventDestroy(h);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TVS_OK;
}

tvs_status tvs_get_geometry(const tvs_ctx *c, tvs_geometry *out) {
    if (!c || !out) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    *out = c->geo;
    return TVS_OK;
}

tvs_status tvs_upload_dem_f32(tvs_ctx *c, const float *elev, size_t count) {
    tvs::TraceRange trace_range("tvs_upload_dem_f32");
    if (!c || !elev) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (count != c->geo.dem_size)
        return fail(c, TVS_ERR_SIZE, "elevation count " + std::to_string(count) +
                                         " != width^2 " + std::to_string(c->geo.dem_size));
    // Heights that are all int16 values (bit for bit: -0 is not) are stored
    // as int16, which the batched walk reads two per dword load; the walk
    // widens them back to the same floats, so the results do not change.
    bool narrow = true;
    for (size_t i = 0; i < count && narrow; ++i) {
        const float v = elev[i];
        narrow = v >= -32768.0f && v <= 32767.0f && v == std::floor(v) &&
                 !(v == 0.0f && std::signbit(v));
    }
    if (narrow) {
        std::vector<int16_t> h(count);
        for (size_t i = 0; i < count; ++i) h[i] = static_cast<int16_t>(elev[i]);
        return tvs_upload_dem_i16(c, h.data(), count);
    }
    if (!c->elev_f32) TVS_HIP(c, dev_alloc(&c->elev_f32, tvs::dem_span(count)), "elevations");
    TVS_HIP(c, hipMemcpyAsync(c->elev_f32, elev, count * sizeof(float),
                              hipMemcpyHostToDevice, c->stream),
            "elevation upload");
    if (tvs::dem_copy_offset(count))
        TVS_HIP(c, hipMemcpyAsync(c->elev_f32 + tvs::dem_copy_offset(count), c->elev_f32,
                                  count * sizeof(float), hipMemcpyDeviceToDevice, c->stream),
                "elevation copy");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "elevation upload");
    dev_free(c->elev_i16);
    c->has_i16 = false;
    c->has_h16 = false;
    c->has_f32 = true;
    bool integral = true;
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = 0; i < count && integral; ++i) {
        const float v = elev[i];
        if (!(std::fabs(v) < 16777216.0f) || v != std::floor(v)) integral = false;
        lo = std::min(lo, static_cast<double>(v));
        hi = std::max(hi, static_cast<double>(v));
    }
    const bool spaced_ok = spaced_divide_holds(elev, count, c->params.observer_height);
    return upload_common(c, integral, spaced_ok, integral ? static_cast<int64_t>(lo) : 0,
                         integral ? static_cast<int64_t>(hi) : -1);
}

tvs_status tvs_upload_dem_i16(tvs_ctx *c, const int16_t *elev, size_t count) {
    tvs::TraceRange trace_range("tvs_upload_dem_i16");
    if (!c || !elev) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (count != c->geo.dem_size)
        return fail(c, TVS_ERR_SIZE, "elevation count " + std::to_string(count) +
                                         " != width^2 " + std::to_string(c->geo.dem_size));
    if (!c->elev_i16) TVS_HIP(c, dev_alloc(&c->elev_i16, tvs::dem_span(count)), "elevations");
    if (!c->elev_f32) TVS_HIP(c, dev_alloc(&c->elev_f32, tvs::dem_span(count)), "elevations");
    TVS_HIP(c, hipMemcpyAsync(c->elev_i16, elev, count * sizeof(int16_t),
                              hipMemcpyHostToDevice, c->stream),
            "elevation upload");
    TVS_HIP(c, tvs::launch_widen_i16(c->stream, c->elev_i16, c->elev_f32, count),
            "elevation widen");
    if (tvs::dem_copy_offset(count)) {
        TVS_HIP(c, hipMemcpyAsync(c->elev_i16 + tvs::dem_copy_offset(count), c->elev_i16,
                                  count * sizeof(int16_t), hipMemcpyDeviceToDevice, c->stream),
                "elevation copy");
        TVS_HIP(c, hipMemcpyAsync(c->elev_f32 + tvs::dem_copy_offset(count), c->elev_f32,
                                  count * sizeof(float), hipMemcpyDeviceToDevice, c->stream),
                "elevation copy");
    }
    TVS_HIP(c, hipStreamSynchronize(c->stream), "elevation upload");
    c->has_i16 = true;
    c->has_f32 = true;
    int64_t lo = 32767, hi = -32768;
    for (size_t i = 0; i < count; ++i) {
        lo = std::min<int64_t>(lo, elev[i]);
        hi = std::max<int64_t>(hi, elev[i]);
    }
    // Heights within 4096 of each other also as f16 offsets from an integer
    // C with every |h - C| <= 2048 (f16 holds those integers exactly): C is
    // the value of [hi - 2048, lo + 2048] nearest 0, which keeps C - h0 exact
    // in f32 for the most observers (the walk checks each one).
    c->has_h16 = false;
    if (hi - lo <= 4096) {
        const int64_t off = std::min(std::max<int64_t>(0, hi - 2048), lo + 2048);
        if (!c->elev_h16) TVS_HIP(c, dev_alloc(&c->elev_h16, tvs::dem_span(count)), "elevations");
        TVS_HIP(c, tvs::launch_half_heights(c->stream, c->elev_i16, c->elev_h16,
                                            tvs::dem_span(count), static_cast<int32_t>(off)),
                "half heights");
        TVS_HIP(c, hipStreamSynchronize(c->stream), "half heights");
        c->h16_offset = static_cast<float>(off);
        c->has_h16 = true;
    }
    return upload_common(c, true, false, lo, hi);
}

tvs_status tvs_compute_bands(tvs_ctx *c, uint32_t first, uint32_t count) {
    tvs::TraceRange trace_range("tvs_compute_bands");
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (first >= tvs::kSectorSteps || count > tvs::kSectorSteps - first)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "sector range outside 0..180");
    if (count == 0) return TVS_OK;
    tvs_status st = check_band_geometry(c);
    if (st != TVS_OK) return st;
    TVS_HIP(c, hipEventRecord(c->ev[4], c->stream), "event");
    std::vector<uint32_t> sorted_path;
    if (c->band_path == TVS_BANDS_SELECT) {
        TVS_HIP(c, tvs::launch_band_select(c, first, count, c->d_axes, c->d_select_status),
                "band selection");
        std::vector<uint32_t> status(count, 0);
        TVS_HIP(c, hipMemcpyAsync(status.data(), c->d_select_status + first,
                                  count * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream),
                "band selection status");
        TVS_HIP(c, hipStreamSynchronize(c->stream), "band selection");
        for (uint32_t i = 0; i < count; ++i)
            if (status[i]) sorted_path.push_back(first + i);
    } else {
        for (uint32_t s = first; s < first + count; ++s) sorted_path.push_back(s);
    }
    if (!sorted_path.empty()) {
        if ((st = ensure_sort_scratch(c)) != TVS_OK) return st;
        for (uint32_t s : sorted_path) {
            TVS_HIP(c, tvs::launch_axes(c, s, false), "axes sort");
            TVS_HIP(c, tvs::launch_band_compile(c, s), "band compile");
        }
    }
    c->bands_sorted_sectors = static_cast<uint32_t>(sorted_path.size());
    TVS_HIP(c, tvs::launch_step_tables(c, first, count), "step tables");
    std::fill(c->short_div_checked.begin() + first, c->short_div_checked.begin() + first + count, 0);
    TVS_HIP(c, hipEventRecord(c->ev[5], c->stream), "event");
    TVS_HIP(c, hipEventSynchronize(c->ev[5]), "band computation");
    float ms = 0.0f;
    TVS_HIP(c, hipEventElapsedTime(&ms, c->ev[4], c->ev[5]), "event time");
    c->timing.bands_ms = ms;
    return read_meta(c, first, count);
}

tvs_status tvs_set_short_divide(tvs_ctx *c, uint32_t enable) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    c->short_div = enable != 0;
    return TVS_OK;
}

tvs_status tvs_set_half_heights(tvs_ctx *c, uint32_t enable) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    c->half_heights = enable != 0;
    return TVS_OK;
}

tvs_status tvs_get_half_heights(const tvs_ctx *c, uint32_t *enabled, uint32_t *available) {
    if (!c || !enabled || !available) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    *enabled = c->half_heights ? 1u : 0u;
    *available = c->has_i16 && c->has_h16 ? 1u : 0u;
    return TVS_OK;
}

tvs_status tvs_get_fast_divide(const tvs_ctx *c, uint32_t *holds) {
    if (!c || !holds) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    *holds = c->fast_div_ok ? 1u : 0u;
    return TVS_OK;
}

tvs_status tvs_get_short_divide(tvs_ctx *c, uint32_t *enabled, uint32_t *verified_sectors) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (enabled) *enabled = c->short_div ? 1u : 0u;
    if (verified_sectors) {
        uint32_t n = 0;
        if (c->short_div && c->fast_div_ok && c->div_domain) {
            std::vector<uint32_t> bad(tvs::kSectorSteps, 1u);
            TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
            TVS_HIP(c, hipMemcpy(bad.data(), c->short_div_bad, bad.size() * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost),
                    "divide check readback");
            for (uint32_t s = 0; s < tvs::kSectorSteps; ++s)
                if (c->short_div_checked[s] && bad[s] == 0u) ++n;
        }
        *verified_sectors = n;
    }
    return TVS_OK;
}

tvs_status tvs_set_band_method(tvs_ctx *c, tvs_band_method m) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (m != TVS_BANDS_SELECT && m != TVS_BANDS_SORT)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "unknown band method");
    c->band_path = m;
    return TVS_OK;
}

tvs_status tvs_bands_sorted_count(const tvs_ctx *c, uint32_t *out) {
    if (!c || !out) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null argument");
    *out = c->bands_sorted_sectors;
    return TVS_OK;
}

tvs_status tvs_compute_axes(tvs_ctx *c, uint32_t sector, float *distances,
                            uint32_t *sight_ordered_map, uint32_t *sector_ordered) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (sector >= tvs::kSectorSteps)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "sector outside 0..180");
    tvs_status st = ensure_sort_scratch(c);
    if (st != TVS_OK) return st;
    TVS_HIP(c, tvs::launch_axes(c, sector, true), "axes sort");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "axes sort");
    const size_t n = c->geo.dem_size;
    if (distances)
        TVS_HIP(c, hipMemcpy(distances, c->sort.dist_f32, n * sizeof(float), hipMemcpyDeviceToHost),
                "axes readback");
    if (sight_ordered_map)
        TVS_HIP(c, hipMemcpy(sight_ordered_map, c->sort.sight_map, n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost),
                "axes readback");
    if (sector_ordered)
        TVS_HIP(c, hipMemcpy(sector_ordered, c->sort.id_a, n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost),
                "axes readback");
    return TVS_OK;
}

static tvs_status set_band_slot(tvs_ctx *c, uint32_t slot, const int32_t *deltas,
                                const float *dists, size_t n) {
    if (!deltas || !dists) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null band data");
    if (n != c->geo.band_deltas_size)
        return fail(c, TVS_ERR_SIZE, "band data length " + std::to_string(n) +
                                         " != band_size - 1 = " +
                                         std::to_string(c->geo.band_deltas_size));
    const size_t tb = static_cast<size_t>(slot) * c->tab.stride;
    if (n) {
        TVS_HIP(c, hipMemcpyAsync(c->tab.delta + tb, deltas, n * sizeof(int32_t),
                                  hipMemcpyHostToDevice, c->stream),
                "band upload");
        TVS_HIP(c, hipMemcpyAsync(c->tab.dist + tb, dists, n * sizeof(float),
                                  hipMemcpyHostToDevice, c->stream),
                "band upload");
    }
    TVS_HIP(c, tvs::launch_step_tables(c, slot, 1), "step tables");
    c->short_div_checked[slot] = 0;
    return read_meta(c, slot, 1);
}

tvs_status tvs_set_band(tvs_ctx *c, uint32_t sector, const int32_t *deltas,
                        const float *dists, size_t n) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (sector >= tvs::kSectorSteps)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "sector outside 0..180");
    return set_band_slot(c, sector, deltas, dists, n);
}

tvs_status tvs_get_band(tvs_ctx *c, uint32_t sector, int32_t *deltas, float *dists,
                        size_t n) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (sector >= tvs::kSectorSteps)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "sector outside 0..180");
    if (!c->band_ready[sector]) return fail(c, TVS_ERR_STATE, "sector has no band data");
    if (n != c->geo.band_deltas_size) return fail(c, TVS_ERR_SIZE, "band data length");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    const size_t tb = static_cast<size_t>(sector) * c->tab.stride;
    if (n && deltas)
        TVS_HIP(c, hipMemcpy(deltas, c->tab.delta + tb, n * sizeof(int32_t), hipMemcpyDeviceToHost),
                "band readback");
    if (n && dists)
        TVS_HIP(c, hipMemcpy(dists, c->tab.dist + tb, n * sizeof(float), hipMemcpyDeviceToHost),
                "band readback");
    return TVS_OK;
}

static tvs_status begin_capture(tvs_ctx *c, uint32_t capture) {
    TVS_HIP(c, hipMemsetAsync(c->overflow, 0, sizeof(unsigned long long), c->stream),
            "overflow reset");
    c->ring_valid = 0;
    if (capture == 0) return TVS_OK;
    tvs_status st = ensure_ring(c, capture);
    if (st != TVS_OK) return st;
    const size_t per = 2ull * c->geo.computable_points_count * c->rings_per_band;
    TVS_HIP(c, tvs::launch_fill_u32(c->stream, c->ring, per * capture, 0u), "ring clear");
    return TVS_OK;
}

}  // extern "C"

namespace tvs {

tvs_status run_sectors(tvs_ctx *c, uint32_t first, uint32_t count, tvs_method method,
                       bool keep_rings) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (first >= tvs::kSectorSteps || count > tvs::kSectorSteps - first)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "sector range outside 0..180");
    for (uint32_t s = first; s < first + count; ++s)
        if (!c->band_ready[s])
            return fail(c, TVS_ERR_STATE, "sector " + std::to_string(s) + " has no band data");
    return run_slots(c, first, count, method, keep_rings);
}

tvs_status run_slots(tvs_ctx *c, uint32_t first, uint32_t count, tvs_method method,
                     bool keep_rings) {
    if (method != TVS_METHOD_PLAIN && method != TVS_METHOD_BATCHED)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "unknown method");
    if (!c->has_f32) return fail(c, TVS_ERR_STATE, "no elevations uploaded");
    const uint32_t N = c->geo.computable_points_count;
    const uint32_t capture = keep_rings ? 0u : std::min(c->params.ring_capture_sectors, count);
    tvs_status st = keep_rings ? TVS_OK : begin_capture(c, capture);
    if (st != TVS_OK) return st;
    c->timing = tvs_timing{0.0f, 0.0f, c->timing.bands_ms, 0u,
                           static_cast<uint64_t>(count) * 2ull * N * c->geo.band_deltas_size};
    if (count == 0 || N == 0) return TVS_OK;
    const size_t per = 2ull * N * c->rings_per_band;
    if (c->walk_hist.empty()) {
        c->walk_hist.assign(2u * tvs_ctx::kWalkHistory, nullptr);
        for (hipEvent_t &h : c->walk_hist) {
            const hipError_t e = hipEventCreate(&h);
            if (e != hipSuccess) {
                for (hipEvent_t &g : c->walk_hist)
                    if (g) (void)hipEventDestroy(g);
                c->walk_hist.clear();
                return fail(c, TVS_ERR_HIP, std::string("walk history events: ") +
                                                hipGetErrorString(e));
            }
        }
    }
    hipEvent_t *hist = &c->walk_hist[2u * (c->walk_runs % tvs_ctx::kWalkHistory)];
    if (method == TVS_METHOD_PLAIN) {
        TVS_HIP(c, hipEventRecord(c->ev[0], c->stream), "event");
        TVS_HIP(c, hipEventRecord(hist[0], c->stream), "event");
        for (uint32_t r = 0; r < count; ++r) {
            uint32_t *ring = r < capture ? c->ring + per * r : nullptr;
            TVS_HIP(c, tvs::launch_walk_plain(c, first + r, ring), "plain walk");
        }
        TVS_HIP(c, hipEventRecord(c->ev[1], c->stream), "event");
        TVS_HIP(c, hipEventRecord(hist[1], c->stream), "event");
        TVS_HIP(c, hipEventRecord(c->ev[2], c->stream), "event");
        c->timing.walk_launches = count;
    } else {
        // Ring capture runs the capture shape; its items must be counted for it.
        const tvs::WalkConfig run_cfg = tvs::effective_walk_config(c);
        const tvs::WalkConfig cfg = capture ? tvs::capture_walk_config(run_cfg) : run_cfg;
        tvs::TileDesc d;
        if ((st = tvs::prepare_tile(c, first, count, capture, cfg, &d)) != TVS_OK)
            return st;
        if ((st = tvs::ensure_short_div(c, c->tab, first, count)) != TVS_OK) return st;
        const uint32_t n_items = d.blocks_per_pass * 2u * count;
        const uint32_t begin[4] = {0u, n_items, 0u, N};
        if (!c->desc_cached || std::memcmp(&d, &c->desc_on_device, sizeof d) != 0 ||
            std::memcmp(begin, c->begin_on_device, sizeof begin) != 0) {
            c->desc_cached = false;
            TVS_HIP(c, hipMemcpyAsync(c->d_tile, &d, sizeof d, hipMemcpyHostToDevice,
                                      c->stream),
                    "tile descriptor");
            TVS_HIP(c, hipMemcpyAsync(c->d_begin, begin, sizeof begin, hipMemcpyHostToDevice,
                                      c->stream),
                    "tile descriptor");
            c->desc_on_device = d;
            std::memcpy(c->begin_on_device, begin, sizeof begin);
            c->desc_cached = true;
        }
        TVS_HIP(c, hipEventRecord(c->ev[0], c->stream), "event");
        TVS_HIP(c, hipEventRecord(hist[0], c->stream), "event");
        TVS_HIP(c, tvs::launch_walk_tiles(c->stream, c->d_tile, c->d_begin, 1, n_items,
                                          c->has_i16, capture > 0, d.fast == 0u, cfg,
                                          d.blocks_per_pass),
                "batched walk");
        TVS_HIP(c, hipEventRecord(c->ev[1], c->stream), "event");
        TVS_HIP(c, hipEventRecord(hist[1], c->stream), "event");
        TVS_HIP(c, tvs::launch_reduce_tiles(c->stream, c->d_tile, c->d_begin + 2, 1, N),
                "sector reduction");
        TVS_HIP(c, hipEventRecord(c->ev[2], c->stream), "event");
        c->timing.walk_launches = 1;
    }
    ++c->walk_runs;
    if (!keep_rings) c->ring_valid = capture;
    return TVS_OK;
}

}  // namespace tvs

extern "C" {

tvs_status tvs_run_sectors(tvs_ctx *c, uint32_t first, uint32_t count, tvs_method method) {
    tvs::TraceRange trace_range("tvs_run_sectors");
    return tvs::run_sectors(c, first, count, method, false);
}

tvs_status tvs_run_sector_with_band(tvs_ctx *c, const int32_t *deltas, const float *dists,
                                    size_t n, float *cum_out, size_t n_out) {
    tvs::TraceRange trace_range("tvs_run_sector_with_band");
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (!c->has_f32) return fail(c, TVS_ERR_STATE, "no elevations uploaded");
    const uint32_t N = c->geo.computable_points_count;
    if (cum_out && n_out != N) return fail(c, TVS_ERR_SIZE, "surface buffer length");
    tvs_status st = set_band_slot(c, tvs::kScratchSector, deltas, dists, n);
    if (st != TVS_OK) return st;
    const uint32_t capture = std::min(c->params.ring_capture_sectors, 1u);
    if (capture == 0) {
        // Without ring capture the batched walk (bit-identical to the plain
        // one) walks the slot, with the four-operation divide: the band
        // changes with every call, and checking it for the short divide
        // would cost more than the two operations it saves on one sector.
        // 1201 x 1201, host to host per call: 0.18-0.20 ms, against
        // 0.24-0.27 ms with the plain walk.
        TVS_HIP(c, tvs::launch_fill_u32(c->stream, c->short_div_bad + tvs::kScratchSector, 1, 1u),
                "divide flag");
        c->short_div_checked[tvs::kScratchSector] = 1;
        if ((st = tvs::run_slots(c, tvs::kScratchSector, 1, TVS_METHOD_BATCHED, false)) != TVS_OK)
            return st;
        if (cum_out) return tvs_read_surfaces(c, cum_out, n_out);
        return TVS_OK;
    }
    if ((st = begin_capture(c, capture)) != TVS_OK) return st;
    if (N) {
        TVS_HIP(c, hipEventRecord(c->ev[0], c->stream), "event");
        TVS_HIP(c, tvs::launch_walk_plain(c, tvs::kScratchSector, capture ? c->ring : nullptr),
                "plain walk");
        TVS_HIP(c, hipEventRecord(c->ev[1], c->stream), "event");
        TVS_HIP(c, hipEventRecord(c->ev[2], c->stream), "event");
    }
    c->ring_valid = capture;
    c->timing = tvs_timing{0.0f, 0.0f, c->timing.bands_ms, N ? 1u : 0u,
                           2ull * N * c->geo.band_deltas_size};
    if (cum_out) return tvs_read_surfaces(c, cum_out, n_out);
    return TVS_OK;
}

// One sector's ring events in compact form: a counting walk, offsets scanned
// on the host, then a writing walk.  Surfaces are not touched (the band
// surfaces of these walks are not reduced).
tvs_status tvs_compact_rings(tvs_ctx *c, uint32_t sector, uint32_t *counts, size_t n_counts,
                             uint32_t *events, size_t events_capacity, uint64_t *n_events) {
    tvs::TraceRange trace_range("tvs_compact_rings");
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    if (sector >= tvs::kSectorSteps)
        return fail(c, TVS_ERR_INVALID_ARGUMENT, "sector outside 0..180");
    if (!c->has_f32) return fail(c, TVS_ERR_STATE, "no elevations uploaded");
    if (!c->band_ready[sector]) return fail(c, TVS_ERR_STATE, "sector has no band data");
    const uint32_t N = c->geo.computable_points_count;
    const size_t bands = 2ull * N;
    if (counts && n_counts != bands) return fail(c, TVS_ERR_SIZE, "counts length != 2N");
    if (N == 0) {
        if (n_events) *n_events = 0;
        return TVS_OK;
    }
    if (!c->d_ring_counts) {
        TVS_HIP(c, dev_alloc(&c->d_ring_counts, bands), "ring counts");
        TVS_HIP(c, dev_alloc(&c->d_ring_offsets,
                             tvs::ring_offsets_len(static_cast<uint32_t>(bands))),
                "ring offsets");
    }
    const tvs::WalkConfig cfg = tvs::capture_walk_config(c->walk_cfg);
    tvs::TileDesc d;
    tvs_status st = tvs::prepare_tile(c, sector, 1, 1, cfg, &d);
    if (st != TVS_OK) return st;
    const uint32_t n_items = d.blocks_per_pass * 2u;
    const uint32_t begin[4] = {0u, n_items, 0u, N};
    d.ring_mode = tvs::kRingCount;
    d.ring_counts = c->d_ring_counts;
    c->desc_cached = false;
    TVS_HIP(c, hipMemcpyAsync(c->d_begin, begin, sizeof begin, hipMemcpyHostToDevice, c->stream), "descriptor");
    // Counts and offsets of this sector are kept until the next upload or
    // band change, so the usual pair of calls (size query, then events)
    // walks the sector's bands for counting once.
    const bool counted = c->ring_cache_gen == c->state_gen && c->ring_cache_sector == sector;
    if (!counted) {
        c->ring_cache_gen = ~uint64_t(0);
        TVS_HIP(c, hipMemcpyAsync(c->d_tile, &d, sizeof d, hipMemcpyHostToDevice, c->stream), "descriptor");
        TVS_HIP(c, tvs::launch_walk_tiles(c->stream, c->d_tile, c->d_begin, 1, n_items,
                                          c->has_i16, true, d.fast == 0u, cfg,
                                          d.blocks_per_pass),
                "ring count walk");
        TVS_HIP(c, tvs::launch_ring_offsets(c->stream, c->d_ring_counts,
                                            static_cast<uint32_t>(bands), c->d_ring_offsets),
                "ring offsets");
        TVS_HIP(c, hipMemcpyAsync(&c->ring_cache_total, c->d_ring_offsets + bands,
                                  sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream),
                "ring total");
    }
    if (counts)
        TVS_HIP(c, hipMemcpyAsync(counts, c->d_ring_counts, bands * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream), "ring counts");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "ring counts");
    if (!counted) {
        c->ring_cache_gen = c->state_gen;
        c->ring_cache_sector = sector;
    }
    const uint64_t total = c->ring_cache_total;
    if (n_events) *n_events = total;
    if (!events) return TVS_OK;
    if (events_capacity < total)
        return fail(c, TVS_ERR_SIZE, "events buffer holds " + std::to_string(events_capacity) +
                                         ", sector has " + std::to_string(total));
    if (c->ring_events_cap < total) {
        dev_free(c->d_ring_events);
        c->ring_events_cap = 0;
        TVS_HIP(c, dev_alloc(&c->d_ring_events, total), "ring events");
        c->ring_events_cap = total;
    }
    d.ring_mode = tvs::kRingCompact;
    d.ring_offsets = c->d_ring_offsets;
    d.ring_events = c->d_ring_events;
    TVS_HIP(c, hipMemcpyAsync(c->d_tile, &d, sizeof d, hipMemcpyHostToDevice, c->stream), "descriptor");
    TVS_HIP(c, tvs::launch_walk_tiles(c->stream, c->d_tile, c->d_begin, 1, n_items, c->has_i16,
                                      true, d.fast == 0u, cfg, d.blocks_per_pass),
            "ring write walk");
    if (total)
        TVS_HIP(c, hipMemcpyAsync(events, c->d_ring_events, total * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream), "ring events");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "ring events");
    return TVS_OK;
}

tvs_status tvs_reset_surfaces(tvs_ctx *c) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    const size_t N = std::max(c->geo.computable_points_count, 1u);
    TVS_HIP(c, hipMemsetAsync(c->cum, 0, N * sizeof(float), c->stream), "reset");
    TVS_HIP(c, hipMemsetAsync(c->cum64, 0, N * sizeof(double), c->stream), "reset");
    return TVS_OK;
}

tvs_status tvs_read_surfaces(tvs_ctx *c, float *out, size_t n) {
    tvs::TraceRange trace_range("tvs_read_surfaces");
    if (!c || !out) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (n != c->geo.computable_points_count) return fail(c, TVS_ERR_SIZE, "surface length");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    if (n) TVS_HIP(c, hipMemcpy(out, c->cum, n * sizeof(float), hipMemcpyDeviceToHost), "readback");
    return TVS_OK;
}

tvs_status tvs_read_surfaces_f64(tvs_ctx *c, double *out, size_t n) {
    if (!c || !out) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (n != c->geo.computable_points_count) return fail(c, TVS_ERR_SIZE, "surface length");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    if (n) TVS_HIP(c, hipMemcpy(out, c->cum64, n * sizeof(double), hipMemcpyDeviceToHost), "readback");
    return TVS_OK;
}

tvs_status tvs_read_rings(tvs_ctx *c, uint32_t slot, uint32_t *out, size_t n) {
    if (!c || !out) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (slot >= c->ring_valid) return fail(c, TVS_ERR_STATE, "ring slot not captured by the last run");
    const size_t per = 2ull * c->geo.computable_points_count * c->rings_per_band;
    if (n != per) return fail(c, TVS_ERR_SIZE, "ring buffer length != 2N*R");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    if (n) TVS_HIP(c, hipMemcpy(out, c->ring + per * slot, n * sizeof(uint32_t), hipMemcpyDeviceToHost),
                   "ring readback");
    return TVS_OK;
}

tvs_status tvs_ring_overflow_count(tvs_ctx *c, uint64_t *out) {
    if (!c || !out) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    unsigned long long v = 0;
    TVS_HIP(c, hipMemcpy(&v, c->overflow, sizeof v, hipMemcpyDeviceToHost), "readback");
    *out = v;
    return TVS_OK;
}

tvs_status tvs_sync(tvs_ctx *c) {
    if (!c) return fail(nullptr, TVS_ERR_INVALID_ARGUMENT, "null context");
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    return TVS_OK;
}

void *tvs_host_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();   // leave no sticky error for the next call
        return nullptr;
    }
    return p;
}

tvs_status tvs_host_free(void *p) {
    if (!p) return TVS_OK;
    return hipHostFree(p) == hipSuccess ? TVS_OK : TVS_ERR_HIP;
}

tvs_status tvs_get_timing(tvs_ctx *c, tvs_timing *out) {
    if (!c || !out) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    if (c->timing.walk_launches) {
        TVS_HIP(c, hipEventSynchronize(c->ev[2]), "event");
        float ms = 0.0f;
        TVS_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]), "event time");
        c->timing.walk_ms = ms;
        TVS_HIP(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]), "event time");
        c->timing.reduce_ms = ms;
    }
    *out = c->timing;
    return TVS_OK;
}

tvs_status tvs_get_walk_history(tvs_ctx *c, float *ms, size_t n, size_t *got) {
    if (!c || (!ms && n)) return fail(c, TVS_ERR_INVALID_ARGUMENT, "null argument");
    const uint64_t kept = std::min<uint64_t>(c->walk_runs, tvs_ctx::kWalkHistory);
    const size_t k = static_cast<size_t>(std::min<uint64_t>(n, kept));
    TVS_HIP(c, hipStreamSynchronize(c->stream), "sync");
    for (size_t j = 0; j < k; ++j) {
        const uint64_t run = c->walk_runs - k + j;
        const hipEvent_t *h = &c->walk_hist[2u * (run % tvs_ctx::kWalkHistory)];
        TVS_HIP(c, hipEventElapsedTime(&ms[j], h[0], h[1]), "event time");
    }
    if (got) *got = k;
    return TVS_OK;
}

}  // extern "C"
