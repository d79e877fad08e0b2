// engine.hpp -- internal declarations shared by the engine's translation
// units (device buffers, per-sector step tables, kernel launchers).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tvs.h"

namespace tvs {

// A named range in rocprofv3's marker trace (`--marker-trace`) around the
// host side of one entry point: upload, band data, a run's enqueue, an
// export.  A library call and nothing else when no profiler is attached.
class TraceRange {
  public:
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

// kernel.rs:13, 17, 21
constexpr float kMaxAngle = -2000.0f;
constexpr float kTanOneRad = 0.0174533f;
constexpr float kEarthRadiusSquared = 12742000.0f;
// axes.rs:35, 47
constexpr uint32_t kSectorSteps = 180;
constexpr float kSectorShift = 0.001f;
// One extra band-table slot holds band data handed in by
// tvs_run_sector_with_band.
constexpr uint32_t kScratchSector = kSectorSteps;
constexpr uint32_t kBandSlots = kSectorSteps + 1;

// Device DEM layout: the n = w*w heights (copy A, elements [0, n)) and,
// for DEMs of up to kDemCopyMaxPoints points, a second copy of them
// starting at the odd element dem_copy_offset(n) (copy B; offset 0: no
// copy).  Two adjacent int16 heights at an odd element index straddle a
// dword in copy A and fill one aligned dword in copy B; the step tables
// send each step whose two-point loads would be misaligned in A to B.
// Both the int16 and the f32 buffers have the layout, so the tables do not
// depend on the uploaded height type.  Measured: 1201^2 and the 64 tiles
// 4 % faster with the copy; 3601^2 3 % slower while every step ran the
// masked update, 4 % faster in the walk's directional item order once the
// update is skipped where no lane sees the point (34.0 against 35.5 ms;
// the XCD-contiguous order 2 % slower, so the tuner picks).  The limit
// keeps every step's byte position, copy included, below 2^32 with room
// (at most ~4 n int16 bytes, twice that for f32 heights).
constexpr uint64_t kDemCopyMaxPoints = uint64_t(1) << 24;
inline uint64_t dem_copy_offset(uint64_t n) { return n <= kDemCopyMaxPoints ? (n | 1u) : 0u; }
inline uint64_t dem_span(uint64_t n) { return dem_copy_offset(n) + n; }
// Parity of the DEM index of every two-point pair's first point: pairs
// start at even columns, and on odd-width DEMs odd rows start at column 1,
// which makes the parity 0; on even-width DEMs it is that of L.
//
// A raw dword buffer load runs at the texture unit's misaligned rate (about
// twice the time) whenever its vector offset or its scalar offset is 2 mod
// 4, even if their sum is aligned (measured on gfx950: 3.6 against 6.7 ns
// per wave64 load per CU).  So the walks keep both aligned: a lane's
// vector offset is that of element pov - pair_parity (TileDesc::pov_shift)
// and every step-table position carries the other element, and each
// sector's bias (the step offsets' lower bound, which places the buffer's
// base below the DEM) is even.
inline uint32_t pair_parity(uint32_t w, uint32_t L) { return (w & 1u) ? 0u : (L & 1u); }

// Steps of one sector's walk are split in four ranges, each a uniform
// (whole-grid) property of the sector:
//   [0, head)              leading irregular steps (d outside the regular
//                          range, e.g. a first band point at distance +0):
//                          IEEE divide; head is 0 when step 0 is regular
//   [head, n_fast)         regular steps: d normal, the fast divide holds
//   [n_fast, tail_start)   anything else: per-step choice of divide
//   [tail_start, n_steps)  the zero tail of band_of_sight.rs:67-68 (delta 0,
//                          distance +0): the point and the angle repeat.
struct SectorMeta {
    int32_t min_off;     // minimum cumulative offset over the steps
    int32_t max_off;     // maximum cumulative offset over the steps
    uint32_t n_fast;
    uint32_t tail_start;
    uint32_t status;     // non-zero: band compile hit a reference panic site
    uint32_t bias;       // max |cumulative offset|: positions below are bias +/- off
    uint32_t head;
    uint32_t pad;
};

// Device-resident per-sector step tables (structure of arrays, `stride`
// entries per sector slot).  Derived from the reference's band data
// (deltas, distances) only; the per-step constants are shared by all 2N
// bands of the sector.
struct StepTables {
    int32_t *delta = nullptr;  // band_deltas            (band_of_sight.rs:165)
    float *dist = nullptr;     // band_distances         (band_of_sight.rs:168)
    int32_t *off = nullptr;    // inclusive prefix sum of delta
    uint32_t *pos_f = nullptr; // 2 (meta.bias + off): forward position, int16 bytes
    uint32_t *pos_b = nullptr; // 2 (meta.bias - off): backward position, int16 bytes
    float *rcp = nullptr;      // RN(1 / dist)
    float *rcp_lo = nullptr;   // RN((1 - dist * rcp) * rcp): 1/dist - rcp
    float *curv = nullptr;     // RN(dist / 12742000)    (kernel.rs:175)
    float *wt = nullptr;       // RN(dist * 0.0174533)   (kernel.rs:216)
    SectorMeta *meta = nullptr;
    uint32_t stride = 0;
};

// One DEM tile of a walk launch: where its heights, step tables and outputs
// live and the geometry the walk needs.  A launch walks `count` sector slots
// from `first_slot`, both directions, for all N points of every tile.
struct TileDesc {
    const void *elev;            // int16 or f32, w*w row-major
    const int32_t *off;          // step tables, slot 0 (see StepTables)
    const uint32_t *pos_f;
    const uint32_t *pos_b;
    const float *dist;
    const float *rcp;
    const float *rcp_lo;
    const float *curv;
    const float *wt;
    const SectorMeta *meta;
    float *partial;              // f32[count][2][N] band surfaces
    float *cum;                  // f32[N] cumulative surfaces
    double *cum64;               // f64[N]
    uint32_t *ring;              // u32[capture_slots][2N][R] or null
    unsigned long long *overflow;
    uint32_t stride, n_steps, first_slot, count;
    uint32_t N, T, L, w;
    uint32_t span;               // elements from elev to the end of copy B
    uint32_t copy_offset;        // first element of copy B
    uint32_t pov_shift;          // pair_parity(w, L): a lane's vector offset is that of
                                 // element pov - pov_shift, the step tables add it back
    float obs;
    uint32_t blocks_per_pass;    // ceil(N / block)
    uint32_t R, capture_slots;
    uint32_t fast;               // heights passed the fast-divide range check
    const uint32_t *short_div_bad;    // per slot: 0 = the short divide was verified for
                                 // every height difference of this DEM (null: off)
    uint32_t ring_mode;          // kRingRecords / kRingCount / kRingCompact
    uint32_t *ring_counts;       // u32[2N]: ring_id per band (kRingCount)
    const uint64_t *ring_offsets;  // u64[2N]: first event of each band (kRingCompact)
    uint32_t *ring_events;       // packed event ids (kRingCompact)
    const uint32_t *order;       // per sector slot, blocks_per_pass block indices in walk
                                 // order (null: ascending)
    // int16 DEMs whose heights all lie within h16_offset +/- 2048: the same
    // layout (both copies) as f16 values h - h16_offset, all exact (null:
    // none).  The two-point walk decodes a height and subtracts the
    // observer's in one mixed-precision FMA from these.
    const void *elev_h16;
    float h16_offset;
};

// Ring capture modes of the tile walk.
constexpr uint32_t kRingRecords = 0;
constexpr uint32_t kRingCount = 1;
constexpr uint32_t kRingCompact = 2;

// Launch shape of the tile walk: points per work item (= threads per
// block), steps unrolled per chunk, XCD-contiguous item mapping.
struct WalkConfig {
    uint32_t block = 256;
    uint32_t unroll = 8;
    bool xcd_contiguous = true;
    bool pipeline = false;    // next chunk's heights requested before this chunk's math
    bool persistent = false;  // a few blocks per CU stride over all items
    bool pair = false;        // one lane walks both bands (forward, backward) of a point
    // Points per lane: 0 one; 1 two adjacent points of a row; 2 two such
    // pairs from two adjacent rows (with `pair`: eight bands per lane).
    uint32_t twin = 0;
    // With xcd_contiguous: the work items of each sector taken in the order
    // of their position across the sector's band direction, so that the
    // items in flight on one XCD form a strip along that direction and
    // their bands share the DEM rows and columns its L2 holds.
    bool directional = false;
};

// tvs_walk_shape::xcd_contiguous of a shape: 0 round-robin items, 1
// XCD-contiguous, 2 XCD-contiguous in directional order.
inline uint32_t xcd_mode(const WalkConfig &cfg) {
    return cfg.directional ? 2u : cfg.xcd_contiguous ? 1u : 0u;
}

// The shape used until a caller sets one or runs the tuner: the fastest
// measured on the 1201 x 1201 workload for each height type (int16: 64
// points per item, 8 steps per chunk, two points and both directions per
// lane; f32: 128 x 8, both directions, one point per lane, where two would
// take two loads).
inline WalkConfig default_walk_config(bool i16, bool xcd_contiguous) {
    return i16 ? WalkConfig{64, 8, xcd_contiguous, false, false, true, true}
               : WalkConfig{128, 8, xcd_contiguous, false, false, true, false};
}

// Untuned single DEMs whose heights (one copy, as stored: int16 or f32)
// take more than this many bytes walk in the directional item order: 2^23
// points with int16 heights, 2^22 with f32.  Measured per pass, same box,
// XCD-contiguous against directional.  int16, 64 x 8 two-point shape:
// 1201^2 to 2400^2 within +-2 %; 3072^2 30.6 against 28.9 ms; 3601^2 35.1
// against 33.6; 4096^2 (heights spanning 2900 m) 108.6 against 68.1.  f32,
// 128 x 8 one-point shape: 1201^2 3.25 against 3.24 ms; 2401^2 28.4
// against 22.9; 3601^2 182.7 against 81.7.  The contiguous runs of whole
// sectors that each XCD takes keep far more of a large DEM in flight than
// its L2 holds.
constexpr uint64_t kDirectionalDefaultBytes = uint64_t(1) << 24;

// The one shape of ring-capturing walks (a checking path): 256 points per
// block, 8 steps per chunk, one band per lane.
inline WalkConfig capture_walk_config(const WalkConfig &cfg) {
    return WalkConfig{256, 8, cfg.xcd_contiguous, false, false, false, false};
}

// Blocks per pass (per sector and direction, or per sector with `pair`) of
// the tile walk for N = T*T points: BLOCK points per block, or with `twin`
// the T*(T/2) column pairs two per lane plus (odd T) the T last-column
// points one per lane.
// Lane groups of the two-point walks: the column pairs of every row (twin
// 1), or of every two adjacent rows (twin 2: rows 2r and 2r + 1, the last
// row alone when T is odd).
__host__ __device__ inline uint32_t twin_lane_groups(uint32_t T, uint32_t twin) {
    return twin == 2u ? ((T + 1u) / 2u) * (T / 2u) : T * (T / 2u);
}

inline uint32_t walk_blocks_per_pass(uint32_t T, const WalkConfig &cfg) {
    if (!cfg.twin) return (T * T + cfg.block - 1) / cfg.block;
    const uint32_t groups = twin_lane_groups(T, cfg.twin);
    return (groups + cfg.block - 1) / cfg.block + ((T & 1u) ? (T + cfg.block - 1) / cfg.block : 0u);
}

// Compiled (block, unroll, pipeline, persistent, pair, twin) shapes of the
// tile walk (twin: points per lane as WalkConfig::twin).
#define TVS_WALK_SHAPES(X)                                                               \
    X(256, 8, false, false, false, false) X(128, 8, false, false, false, false)          \
    X(64, 8, false, false, false, false) X(512, 8, false, false, false, false)           \
    X(256, 4, false, false, false, false) X(256, 16, false, false, false, false)         \
    X(128, 16, false, false, false, false) X(256, 8, true, false, false, false)          \
    X(128, 8, true, false, false, false) X(256, 4, true, false, false, false)            \
    X(512, 8, true, false, false, false) X(256, 8, false, true, false, false)            \
    X(128, 8, false, true, false, false) X(64, 8, false, true, false, false)             \
    X(64, 8, false, false, true, false) X(128, 8, false, false, true, false)             \
    X(256, 8, false, false, true, false) X(64, 4, false, false, true, false)             \
    X(128, 4, false, false, true, false) X(64, 8, false, false, false, true)             \
    X(128, 8, false, false, false, true) X(64, 4, false, false, false, true)             \
    X(64, 4, false, false, true, true) X(128, 4, false, false, true, true)               \
    X(64, 8, false, false, true, true) X(64, 4, false, true, true, true)                 \
    X(64, 8, false, true, false, true) X(64, 4, false, false, true, 2)

// Scratch for the device Axes::compute (keys, ids, radix-sort ping-pong,
// histograms).
struct SortScratch {
    uint32_t *key_a = nullptr, *key_b = nullptr;
    uint32_t *id_a = nullptr, *id_b = nullptr;
    uint32_t *sight_map = nullptr;   // axes.rs:121-124
    uint32_t *sector_ids = nullptr;  // axes.rs:127-129
    uint32_t *hist = nullptr;
    float *dist_f32 = nullptr;       // axes.distances (axes.rs:187-189)
    size_t n = 0;
    uint32_t tiles = 0;
};

}  // namespace tvs

struct tvs_ctx {
    tvs_geometry geo{};
    tvs_params params{};
    uint32_t rings_per_band = 0;   // R
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    // What the last batched run copied into d_tile / d_begin (valid while
    // desc_cached): an identical descriptor is not copied again, which
    // spares back-to-back runs two host-to-device copies each.  Every other
    // writer of d_tile / d_begin clears desc_cached.
    tvs::TileDesc desc_on_device{};
    uint32_t begin_on_device[4] = {};
    bool desc_cached = false;
    // Start and end of the walk of each of the last kWalkHistory runs of
    // tvs_run_sectors (slot = run index mod kWalkHistory), created on the
    // first run.
    static constexpr uint32_t kWalkHistory = 64;
    std::vector<hipEvent_t> walk_hist;
    uint64_t walk_runs = 0;

    float *elev_f32 = nullptr;
    int16_t *elev_i16 = nullptr;
    bool has_f32 = false, has_i16 = false;
    // The heights as f16 offsets from h16_offset (see TileDesc::elev_h16),
    // when the int16 DEM's range allows; half_heights false turns their use
    // off (the walk then converts the int16 heights; results are the same).
    uint16_t *elev_h16 = nullptr;
    bool has_h16 = false;
    bool half_heights = true;
    float h16_offset = 0.0f;
    // Height-difference bounds proven on the host for the fast divide.
    bool fast_div_ok = false;
    float min_abs_dh = 0.0f, max_abs_dh = 0.0f;
    // Short divide: every height difference the DEM can produce (device),
    // per-slot verification results and which slots are verified.
    bool short_div = true;                 // caller's switch (default on)
    float *div_domain = nullptr;
    uint32_t div_domain_n = 0;
    uint32_t *short_div_bad = nullptr;          // u32[kSlots]
    std::vector<uint8_t> short_div_checked;     // per slot

    tvs::StepTables tab;
    std::vector<uint8_t> band_ready;   // per slot: band data resident
    std::vector<tvs::SectorMeta> meta_host;

    tvs::WalkConfig walk_cfg;
    bool walk_cfg_set = false;   // false: default_walk_config for the uploaded heights
    uint32_t *d_ring_counts = nullptr;    // compact ring export
    uint64_t *d_ring_offsets = nullptr;   // u64[2N + 1 + tiles]: offsets, total, scan scratch
    uint32_t *d_ring_events = nullptr;
    uint32_t *d_order = nullptr;          // directional item order (see WalkConfig)
    uint32_t *d_obs = nullptr;            // tvs_observer_walk: counts, surfaces, event slots
    size_t obs_cap = 0;                   // u32 entries of d_obs
    uint64_t order_key = 0;               // (T, w, block, twin) it was built for; 0: none
    size_t ring_events_cap = 0;
    // Changes with every upload and band change; the compact ring export
    // keeps the counts and offsets of the last sector it counted under it.
    uint64_t state_gen = 0;
    uint64_t ring_cache_gen = ~uint64_t(0);
    uint32_t ring_cache_sector = ~0u;
    uint64_t ring_cache_total = 0;
    tvs::TileDesc *d_tile = nullptr;     // device copy of this DEM's descriptor
    uint32_t *d_begin = nullptr;         // {0, items} and {0, N}
    float *cum = nullptr;      // f32[N]  cumulative surfaces
    double *cum64 = nullptr;   // f64[N]
    float *partial = nullptr;  // f32[sectors][2][N] (batched method)
    size_t partial_cap = 0;    // sectors the partial buffer holds
    uint32_t *ring = nullptr;  // u32[capture][2N][R]
    size_t ring_slots = 0;
    uint32_t ring_valid = 0;   // slots filled by the last run
    unsigned long long *overflow = nullptr;

    tvs::SortScratch sort;
    double *d_axes = nullptr;      // per sector: sin/cos of sight and sector axes
    uint32_t *d_select_status = nullptr;
    int band_path = TVS_BANDS_SELECT;
    uint32_t bands_sorted_sectors = 0;   // sectors of the last call that took the sorts
    tvs_timing timing{};
    std::string err;
};

namespace tvs {

// ---- launchers (bands.hip) ----
hipError_t launch_axes(tvs_ctx *c, uint32_t sector, bool want_distances);
hipError_t launch_band_compile(tvs_ctx *c, uint32_t sector);
hipError_t launch_step_tables(tvs_ctx *c, uint32_t first_slot, uint32_t count);
hipError_t launch_band_select(tvs_ctx *c, uint32_t first, uint32_t count,
                              const void *d_axes, uint32_t *d_status);
uint32_t band_select_capacity();

// ---- launchers (walk.hip) ----
hipError_t launch_walk_plain(tvs_ctx *c, uint32_t slot, uint32_t *ring_out);
bool walk_shape_supported(const WalkConfig &cfg);
// The shape a batched run of this context uses.
WalkConfig effective_walk_config(const tvs_ctx *c);
// Per slot in [first, first + count): bad[slot] = 1 if the short divide
// differs from the split-reciprocal divide for any value of `domain` at any
// regular step of the slot's table (bad must be zero on entry).
hipError_t launch_short_div_check(hipStream_t s, const StepTables &tab, uint32_t first,
                             uint32_t count, const float *domain, uint32_t n_domain,
                             uint32_t *bad);
// Every compiled shape the tuners time: both item mappings (and the
// directional order for two points per lane), persistent shapes once.
std::vector<WalkConfig> compiled_walk_shapes();
// Per sector slot in [0, n_slots): the walk order of the launch shape's
// blocks_per_pass blocks for WalkConfig::directional.
std::vector<uint32_t> directional_walk_order(uint32_t T, uint32_t w, const WalkConfig &cfg,
                                             uint32_t n_slots);
// ieee_only: no tile of the launch has the fast divides (TileDesc::fast
// == 0); the kernel then unrolls the IEEE-divide loop.  Either value gives
// the same results.  pass_items: the one tile's blocks_per_pass, which lets
// a directional launch split every pass across the XCDs (0: not known,
// e.g. several tiles; the directional order then runs XCD-contiguous).
hipError_t launch_walk_tiles(hipStream_t s, const TileDesc *tiles,
                             const uint32_t *item_begin, uint32_t n_tiles,
                             uint32_t n_items, bool i16, bool capture, bool ieee_only,
                             const WalkConfig &cfg, uint32_t pass_items = 0);
hipError_t launch_reduce_tiles(hipStream_t s, const TileDesc *tiles,
                               const uint32_t *point_begin, uint32_t n_tiles,
                               uint32_t n_points);
hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v);
// Entries of `offsets` a compact ring export of n bands needs: n offsets,
// the total, and the scan's per-tile scratch.
size_t ring_offsets_len(uint32_t n);
// offsets[k] = sum over j < k of (counts[j] - 1) (each band's events after
// its PoV slot), offsets[n] = the total.
hipError_t launch_ring_offsets(hipStream_t s, const uint32_t *counts, uint32_t n,
                               uint64_t *offsets);
hipError_t launch_widen_i16(hipStream_t s, const int16_t *in, float *out,
                            size_t n);
hipError_t launch_half_heights(hipStream_t s, const int16_t *in, uint16_t *out, size_t n,
                               int32_t offset);

// ---- host side of a run (api.cpp) ----
// tvs_run_sectors; with keep_rings the run records no rings and leaves the
// ring buffer, the overflow counter and the captured-slot count untouched
// (the tuner's timing runs).
tvs_status run_sectors(tvs_ctx *c, uint32_t first, uint32_t count, tvs_method method,
                       bool keep_rings);
// run_sectors on step-table slots [first, first + count) without the sector
// range and band checks (the scratch slot of tvs_run_sector_with_band).
tvs_status run_slots(tvs_ctx *c, uint32_t first, uint32_t count, tvs_method method,
                     bool keep_rings);
tvs_status validate_slots(tvs_ctx *c, uint32_t first, uint32_t count);
tvs_status prepare_tile(tvs_ctx *c, uint32_t first, uint32_t count, uint32_t capture,
                        const WalkConfig &cfg, TileDesc *out, bool allocate = true);
// Verifies the short divide for the unchecked slots of [first, first +
// count) against step tables `tab` (launches only; allocates nothing).
tvs_status ensure_short_div(tvs_ctx *c, const StepTables &tab, uint32_t first, uint32_t count);

// ---- host geometry (geometry.cpp) ----
tvs_status dem_geometry(uint32_t width, float scale, uint32_t los,
                        tvs_geometry *out);
void sector_sin_cos(uint32_t sector, bool sector_axis, double *s, double *c);

}  // namespace tvs
