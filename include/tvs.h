/*
 * tvs.h -- C ABI of the MI355X total-viewshed engine (libtvs_hip.so).
 *
 * Plain C types only: pointers, sizes, scalars.  Every entry point returns a
 * tvs_status and never aborts.  One context is used by one host thread at a
 * time; each context owns one HIP stream and its device buffers.
 *
 * Each declaration names the reference interface it replaces
 * (tombh/total-viewsheds; paths relative to crates/total-viewsheds/src/
 * unless written crates/shader/src/...).
 */
#ifndef TVS_H
#define TVS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TVS_ABI_VERSION 1u

/* Status codes.  The non-zero values map the reference's `bail!` and panic
 * sites onto error returns. */
typedef enum tvs_status {
    TVS_OK = 0,
    TVS_ERR_INVALID_ARGUMENT = 1, /* null pointer, zero width, bad enum       */
    TVS_ERR_LOS_TOO_LONG = 2,     /* dem.rs:56-62 bail!                        */
    TVS_ERR_BAND_TOO_SHORT = 3,   /* band_of_sight.rs:158-168 index panic
                                     (band_size <= 2) or a window past the DEM */
    TVS_ERR_BAND_OUT_OF_DEM = 4,  /* kernel.rs:164 index panic: a band leaves
                                     the elevation array                       */
    TVS_ERR_SIZE = 5,             /* a buffer length does not match the DEM    */
    TVS_ERR_STATE = 6,            /* call order (e.g. run before upload)       */
    TVS_ERR_NO_DEVICE = 7,        /* no HIP device / gpu.rs:84-89 analogue     */
    TVS_ERR_OUT_OF_MEMORY = 8,
    TVS_ERR_HIP = 9,              /* any other HIP runtime error               */
    TVS_ERR_IO = 10               /* file read/write (.bt, band cache)         */
} tvs_status;

/* Walk implementations (the `ComputeType` slot of config.rs:57-64 /
 * compute.rs:188-198 that `Cuda => unimplemented!()` leaves empty). */
typedef enum tvs_method {
    /* One sector per launch, FP32 heights, the reference loop literally
     * (IEEE f32 divide per step), one thread per band. */
    TVS_METHOD_PLAIN = 0,
    /* All requested sectors in one launch: per-step constants hoisted, the
     * divide done as a correctly rounded reciprocal-plus-residual sequence,
     * heights read as int16 when an int16 DEM was uploaded, per-sector band
     * surfaces reduced in sector order (bit-identical f32 TVS) plus an FP64
     * per-point sum. */
    TVS_METHOD_BATCHED = 1
} tvs_method;

/* dem.rs:11-84 (`DEM` scalars) + band_of_sight.rs:54-58, 95, 149-150. */
typedef struct tvs_geometry {
    uint32_t dem_width;               /* w                                  */
    uint32_t dem_size;                /* w*w                                */
    float scale;                      /* metres per point                   */
    uint32_t max_line_of_sight;       /* metres                             */
    uint32_t max_los_as_points;       /* L  (dem.rs:53)                     */
    uint32_t computable_points_count; /* N  (dem.rs:87-94)                  */
    uint32_t tvs_width;               /* T = isqrt(N) (dem.rs:82)           */
    uint32_t band_size;               /* B = L + 1 (dem.rs:78)              */
    uint32_t band_deltas_size;        /* B - 1: steps per band              */
    uint32_t band_samples;            /* floor(2B*sqrt2) (band_of_sight:95) */
    uint32_t real_steps;              /* min(2B - bs/2, B-1): filled prefix */
} tvs_geometry;

/* Context parameters: the arguments of `Compute::new` (compute.rs:27-33)
 * and of `GPU::new` (gpu.rs:55-61) that are not buffers. */
typedef struct tvs_params {
    uint32_t dem_width;
    float scale;
    uint32_t max_line_of_sight;  /* metres; 0 = default (main.rs:36-38)      */
    float rings_per_km;          /* compute.rs:94-98 (reference default 5)  */
    float observer_height;       /* metres (config.rs:30-36; reference 1.8
                                    in Compute::new, compute.rs:41)          */
    uint32_t ring_capture_sectors; /* sectors whose ring records one run
                                      keeps (0 = none, as the SPIR-V build,
                                      kernel.rs:302-309)                    */
} tvs_params;

/* Per-phase device times of the last run / band computation, from HIP
 * events on the context's stream (milliseconds). */
typedef struct tvs_timing {
    float walk_ms;         /* all walk-kernel launches of the last run      */
    float reduce_ms;       /* sector-order reduction (batched method)       */
    float bands_ms;        /* last tvs_compute_bands: keys, sorts, compile  */
    uint32_t walk_launches;
    uint64_t point_steps;  /* sectors * 2N * (B-1) walked by the last run   */
} tvs_timing;

typedef struct tvs_ctx tvs_ctx;

/* ---- geometry (host only, no device needed) --------------------------- */

/* `DEM::new` (dem.rs:44-84).  Returns TVS_ERR_LOS_TOO_LONG as dem.rs:56-62
 * bails. */
tvs_status tvs_dem_geometry(uint32_t dem_width, float scale,
                            uint32_t max_line_of_sight, tvs_geometry *out);

/* Default max line of sight `(w div 3) * scale` (main.rs:36-38). */
uint32_t tvs_default_max_line_of_sight(uint32_t dem_width, double scale);

/* `Compute::ring_count_per_band` (compute.rs:94-98). */
uint32_t tvs_ring_count_per_band(float rings_per_km,
                                 uint32_t max_line_of_sight);

/* ---- context ------------------------------------------------------------ */

/* `GPU::new` (gpu.rs:55-176) + `Compute::new` (compute.rs:27-84): device,
 * stream and the per-DEM constants (kernel.rs:38-58).  Elevations come
 * separately (tvs_upload_dem_*). */
tvs_status tvs_create(const tvs_params *params, tvs_ctx **out);

/* Frees every device buffer and the stream. */
tvs_status tvs_destroy(tvs_ctx *ctx);

/* Geometry of the context's DEM (as tvs_dem_geometry). */
tvs_status tvs_get_geometry(const tvs_ctx *ctx, tvs_geometry *out);

/* Elevation upload, `GPU::setup_buffers` elevations (gpu.rs:195-200), from
 * the f32 array main.rs:51-56 builds.  Copies; the caller keeps ownership.
 * Heights that are all int16 values (no -0) are stored as by
 * tvs_upload_dem_i16: the same floats, read at 2 bytes per height. */
tvs_status tvs_upload_dem_f32(tvs_ctx *ctx, const float *elevations,
                              size_t count);

/* Elevation upload from the `.bt` int16 samples (input.rs:135-171) without
 * the f32 widening.  Also keeps an f32 copy for TVS_METHOD_PLAIN; the
 * widening is exact (main.rs:52-53). */
tvs_status tvs_upload_dem_i16(tvs_ctx *ctx, const int16_t *elevations,
                              size_t count);

/* ---- band data ---------------------------------------------------------- */

/* `DEM::calculate_axes` + `DEM::compile_band_data` (dem.rs:135-139,
 * axes.rs:103-130, band_of_sight.rs:61-173) for sectors
 * [first, first+count), on the device, from f64 rotated-coordinate keys
 * (see tvs_band_method for the two ways of ordering them).  The band data
 * stays resident for the walk.  Synchronous. */
tvs_status tvs_compute_bands(tvs_ctx *ctx, uint32_t first_sector,
                             uint32_t sector_count);

/* The batched walk's divide.  Regular steps normally use a four-operation
 * correctly rounded divide (split reciprocal + exact residual step).  With
 * the short divide on (the default), a sector whose every step gives the
 * same quotient with a two-operation divide, for every height difference
 * the uploaded DEM can produce (checked on the device before the first run
 * after an upload or band change), uses the two-operation one; results
 * are identical either way.  `verified_sectors` counts the sectors that
 * passed the last checks.  No reference counterpart (kernel.rs:175 is a
 * plain f32 divide). */
tvs_status tvs_set_short_divide(tvs_ctx *ctx, uint32_t enable);
tvs_status tvs_get_short_divide(tvs_ctx *ctx, uint32_t *enabled,
                                uint32_t *verified_sectors);
/* The batched walk's height decode.  With half heights on (the default),
 * an int16 DEM whose heights span at most 4096 is also kept as f16 values
 * h - C (C an integer, every |h - C| <= 2048, all exact); the two-point
 * walk then forms each height difference h - h0 in one mixed-precision
 * FMA, h16 + (C - h0), for every wave whose observers all have C - h0
 * exact in f32 (it checks), instead of converting the int16 height and
 * subtracting.  The difference is the same f32 value either way, so results
 * are identical.  `available` = 1 when the uploaded DEM has the f16 form.
 * No reference counterpart (kernel.rs:160-175 subtracts f32 heights). */
tvs_status tvs_set_half_heights(tvs_ctx *ctx, uint32_t enable);
tvs_status tvs_get_half_heights(const tvs_ctx *ctx, uint32_t *enabled,
                                uint32_t *available);
/* `holds` = 1 when the uploaded heights admit the four-operation divide:
 * integer heights whose differences from every observer height stay within
 * the proof's range, or non-integer heights that are all 0 or of magnitude
 * in [2^-17, 2^38], observer heights (height + observer) likewise.  0: the
 * batched walk divides with the IEEE divide (same results, slower).  The
 * short divide applies to integer heights only.  No reference counterpart. */
tvs_status tvs_get_fast_divide(const tvs_ctx *ctx, uint32_t *holds);

/* How tvs_compute_bands finds each sector's window and its sight order.
 * SELECT (default): per-row binary searches over the monotone rows of the
 * rotated key pick the window exactly, and the samples are ordered by
 * (sight key, id) -- the same data the two stable sorts give, without
 * them.  SORT: the two full stable radix sorts of axes.rs:103-130 and the
 * inverse sight map, as tvs_compute_axes runs them.  SELECT falls back to
 * SORT for a sector whose window plus key ties exceeds its buffer. */
typedef enum tvs_band_method {
    TVS_BANDS_SELECT = 0,
    TVS_BANDS_SORT = 1
} tvs_band_method;

tvs_status tvs_set_band_method(tvs_ctx *ctx, tvs_band_method method);

/* Sectors of the last tvs_compute_bands call that went through the sorts. */
tvs_status tvs_bands_sorted_count(const tvs_ctx *ctx, uint32_t *out);

/* Band data supplied from outside (a cache file, another producer): the
 * inputs of `GPU::run` (gpu.rs:387-394), `n` = band_deltas_size. */
tvs_status tvs_set_band(tvs_ctx *ctx, uint32_t sector, const int32_t *deltas,
                        const float *distances, size_t n);

/* Copy a sector's resident band data back (the `dem.band_deltas` /
 * `dem.band_distances` pair of dem.rs:15-17). */
tvs_status tvs_get_band(tvs_ctx *ctx, uint32_t sector, int32_t *deltas,
                        float *distances, size_t n);

/* Run the device `Axes::compute` (axes.rs:103-130) for one sector and copy
 * its outputs back: `distances` (axes.distances, f32[w*w]),
 * `sight_ordered_map` (u32[w*w]), `sector_ordered` (u32[w*w]).  Any output
 * pointer may be NULL. */
tvs_status tvs_compute_axes(tvs_ctx *ctx, uint32_t sector, float *distances,
                            uint32_t *sight_ordered_map,
                            uint32_t *sector_ordered);

/* ---- walk ---------------------------------------------------------------- */

/* `Compute::run`'s sector loop (compute.rs:101-115) over resident band data
 * for sectors [first, first+count), with `compute_sector` dispatching to the
 * chosen method (compute.rs:181-203).  Surfaces accumulate across calls
 * like `cumulative_surfaces` (compute.rs:102, kernel.rs:254) until
 * tvs_reset_surfaces.  Ring records of the first `ring_capture_sectors`
 * sectors of this call are kept (test-mode layout, compute.rs:46-47).
 * Asynchronous on the context's stream: read results or tvs_sync. */
tvs_status tvs_run_sectors(tvs_ctx *ctx, uint32_t first_sector,
                           uint32_t sector_count, tvs_method method);

/* tvs_run_sectors with TVS_METHOD_BATCHED, allocating nothing on the
 * device.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (serving loops, code that must not call hipMalloc).
 * The per-sector band surfaces go to the library's static 64 MiB device
 * scratch instead of a buffer grown on demand, sectors walked in chunks that
 * fit it; results are bit-identical to tvs_run_sectors.  Limits: one
 * sector's band surfaces (2N x 4 bytes) must fit the scratch (N <= 8388608,
 * else TVS_ERR_SIZE); no ring capture (TVS_ERR_INVALID_ARGUMENT when
 * ring_capture_sectors > 0); a directional walk shape runs in the
 * XCD-contiguous order.  Synchronous: returns once the run has finished.
 * Calls on one device share its scratch and are serialised by a lock. */
tvs_status tvs_run_sectors_noalloc(tvs_ctx *ctx, uint32_t first_sector,
                                   uint32_t sector_count);

/* One sector from host band data, returning the cumulative surfaces:
 * exactly `GPU::run` (gpu.rs:387-468) as called by compute_sector_vulkan
 * (compute.rs:217-225).  `cumulative_out` may be NULL.  Synchronous. */
tvs_status tvs_run_sector_with_band(tvs_ctx *ctx, const int32_t *deltas,
                                    const float *distances, size_t n,
                                    float *cumulative_out, size_t n_out);

/* Zero the cumulative surfaces (compute.rs:102). */
tvs_status tvs_reset_surfaces(tvs_ctx *ctx);

/* Cumulative per-point surfaces, f32[N] (the first element of
 * `Compute::run`'s return, compute.rs:114).  Synchronous. */
tvs_status tvs_read_surfaces(tvs_ctx *ctx, float *out, size_t n);

/* The same sums carried in double across sectors (TVS_METHOD_BATCHED keeps
 * them; PLAIN leaves them untouched).  Synchronous. */
tvs_status tvs_read_surfaces_f64(tvs_ctx *ctx, double *out, size_t n);

/* Ring records of captured slot `slot` (the sector first_sector+slot of the
 * last run): 2N*R u32 in the layout kernel.rs:141-243 writes (count, PoV
 * id, then alternating close/open DEM ids), zero elsewhere.  Writes past a
 * band's R slots are dropped and counted instead of spilling into the next
 * band (kernel.rs:141).  Synchronous. */
tvs_status tvs_read_rings(tvs_ctx *ctx, uint32_t slot, uint32_t *out,
                          size_t n);

/* One sector's ring events in compact form, for any number of events per
 * band (no R limit): `counts[k]` = the count the reference stores at slot 0
 * of band k (kernel.rs:243: 1 + events), `events` = every band's events
 * after its PoV slot (kernel.rs:222-239 writes), band after band, k =
 * 0..2N-1 (forward bands, then backward).  Call with events == NULL to get
 * the counts and `*n_events` only.  Does not change the surfaces. */
tvs_status tvs_compact_rings(tvs_ctx *ctx, uint32_t sector, uint32_t *counts,
                             size_t n_counts, uint32_t *events,
                             size_t events_capacity, uint64_t *n_events);

/* One observer's bands: the two kernel ids of observer `tvs_id` (t and
 * t + N, kernel.rs:117-131) in each sector of [first, first+count), walked
 * one wavefront per band (64 consecutive steps per pass, prefix maximum
 * across the lanes, ballots for the ring-event slots), every step as the
 * reference (IEEE divide, zero tail included).  Bands in order (sector,
 * forward then backward): `counts[b]` = the count the reference stores at
 * slot 0 of the band (kernel.rs:243: 1 + events), `band_surfaces[b]` (may
 * be NULL) = the band's f32 surface (kernel.rs:216), `events` = every
 * band's events after its PoV slot, band after band (as tvs_compact_rings
 * packs them); call with events == NULL to get the counts and `*n_events`
 * only.  No counterpart in the reference, which rebuilds one observer's
 * viewshed from the ring data of every band (output/ascii.rs:24-104).
 * Synchronous; surfaces and ring records of runs are not touched. */
tvs_status tvs_observer_walk(tvs_ctx *ctx, uint32_t tvs_id, uint32_t first_sector,
                             uint32_t sector_count, uint32_t *counts,
                             float *band_surfaces, uint32_t *events,
                             size_t events_capacity, uint64_t *n_events);

/* Ring events dropped by the last run because a band had more than R-1. */
tvs_status tvs_ring_overflow_count(tvs_ctx *ctx, uint64_t *out);

/* Wait for the context's stream. */
tvs_status tvs_sync(tvs_ctx *ctx);

/* Page-locked host memory (hipHostMalloc): device-to-host copies into it
 * (the counts and events of tvs_compact_rings and tvs_observer_walk,
 * surfaces, ring records) go straight over the link instead of through the
 * runtime's pageable staging.  NULL when `bytes` is 0 or the allocation
 * fails.  Free with tvs_host_free (NULL is accepted).  Stands for the
 * reference's MAP_READ download buffers (gpu.rs:227-248), which wgpu maps
 * for the host to read. */
void *tvs_host_alloc(size_t bytes);
tvs_status tvs_host_free(void *ptr);

/* Device times of the last run and band computation. */
tvs_status tvs_get_timing(tvs_ctx *ctx, tvs_timing *out);

/* Walk times (ms, as tvs_timing.walk_ms) of the last `n` runs of
 * tvs_run_sectors, oldest first (at most 64 are kept; `*got` says how many
 * were written), so runs issued back to back can be timed without a wait
 * between them.  Waits for the context's stream.  No reference counterpart
 * (measurement). */
tvs_status tvs_get_walk_history(tvs_ctx *ctx, float *ms, size_t n, size_t *got);

/* Message for the last failing call on `ctx` (or on a NULL context). */
const char *tvs_last_error(const tvs_ctx *ctx);

/* Static name of a status code. */
const char *tvs_status_string(tvs_status status);

/* ---- launch shape and autotuning of the batched walk --------------------- */

/* Points per work item (threads per block), steps unrolled per chunk,
 * whether consecutive work items stay on one XCD (`xcd_contiguous`: 0
 * round-robin over the XCDs, 1 contiguous runs of items per XCD, 2 every
 * sector's blocks sorted across its band direction and split in 8
 * contiguous strips, one per XCD; batches of several tiles keep the sorted
 * order with contiguous runs; other values are refused), and whether the next
 * chunk's heights are requested before the current chunk's arithmetic,
 * whether a grid of a few blocks per CU strides over all work items
 * (`persistent`, which ignores `xcd_contiguous`), and whether one lane walks
 * both bands (forward and backward) of its point, sharing the per-step
 * constants (`both_directions`), and whether one lane walks two adjacent
 * points of a row, fetching both int16 heights with one load per step
 * (`two_points`); `ms` is a timing result.
 * No reference counterpart: the Vulkan path's fixed 8x8x4 workgroup
 * (lib.rs:23, gpu.rs:10) is the only shape the reference has. */
typedef struct tvs_walk_shape {
    uint32_t block;
    uint32_t unroll;
    uint32_t xcd_contiguous;
    uint32_t pipeline;
    uint32_t persistent;
    uint32_t both_directions;
    uint32_t two_points;
    float ms;
} tvs_walk_shape;

/* Until a shape is set or tuned, batched runs use the fastest shape measured
 * for the uploaded height type (int16: 64 lanes x 8 steps, two points and
 * both directions per lane; f32: 128 x 8, both directions, one point per
 * lane), with XCD-contiguous items, in the directional order
 * (xcd_contiguous 2) when the heights take more than 16 MiB (int16 DEMs
 * above 2^23 points, f32 above 2^22);
 * tvs_get_walk_shape reports the shape the next run uses. */
tvs_status tvs_set_walk_shape(tvs_ctx *ctx, const tvs_walk_shape *shape);
tvs_status tvs_get_walk_shape(const tvs_ctx *ctx, tvs_walk_shape *out);

/* Number of shapes the tuner times: each compiled shape in both item
 * mappings (two-points-per-lane shapes also in the directional one),
 * persistent shapes once. */
size_t tvs_walk_shape_count(void);

/* Time every compiled shape on the context's own DEM over sectors
 * [first, first+count) (one warm-up run, then the fastest of `reps` timed
 * runs), keep the fastest shape as the context's, return all timings in
 * `results` (up to `max_results`) and the winner in `best`.  Cumulative
 * surfaces, ring records and the ring overflow count are preserved (the
 * timing runs record no rings). */
tvs_status tvs_autotune(tvs_ctx *ctx, uint32_t first_sector, uint32_t sector_count,
                        uint32_t reps, tvs_walk_shape *results, size_t max_results,
                        size_t *n_results, tvs_walk_shape *best);

/* ---- many DEM tiles in one launch ------------------------------------------ */

/* A batch of independent tiles (each a `Compute::new` of its own,
 * compute.rs:27-84, with its own width, spacing and line of sight) whose
 * 180 sectors all run in one walk launch and one reduction launch, replayed
 * from a captured hipGraph.  Tiles of equal (width, band size) share band
 * data (band_of_sight.rs:61-173 reads no elevation). */
typedef struct tvs_batch tvs_batch;

tvs_status tvs_batch_create(const tvs_params *tiles, uint32_t n_tiles, tvs_batch **out);
tvs_status tvs_batch_destroy(tvs_batch *batch);
tvs_status tvs_batch_tile_geometry(const tvs_batch *batch, uint32_t tile,
                                   tvs_geometry *out);
tvs_status tvs_batch_upload_i16(tvs_batch *batch, uint32_t tile,
                                const int16_t *elevations, size_t count);
tvs_status tvs_batch_upload_f32(tvs_batch *batch, uint32_t tile,
                                const float *elevations, size_t count);
/* Band data for all 180 sectors of every distinct (width, band size). */
tvs_status tvs_batch_compute_bands(tvs_batch *batch);
tvs_status tvs_batch_set_walk_shape(tvs_batch *batch, const tvs_walk_shape *shape);
/* The shape the batch's next run uses: the set or tuned one, else the
 * fastest measured for the tiles' height type with items round-robin over
 * the XCDs, in the directional order (xcd_contiguous 2) when a tile's
 * heights take more than 16 MiB. */
tvs_status tvs_batch_get_walk_shape(const tvs_batch *batch, tvs_walk_shape *out);
/* tvs_set_short_divide for every tile. */
tvs_status tvs_batch_set_short_divide(tvs_batch *batch, uint32_t enable);
/* tvs_set_half_heights for every tile. */
tvs_status tvs_batch_set_half_heights(tvs_batch *batch, uint32_t enable);
/* tvs_autotune for a batch: times the walk launch over all tiles in every
 * compiled shape and keeps the fastest.  Only the walk runs, so the tiles'
 * cumulative surfaces are unchanged.  Needs band data and elevations. */
tvs_status tvs_batch_autotune(tvs_batch *batch, uint32_t reps, tvs_walk_shape *results,
                              size_t max_results, size_t *n_results, tvs_walk_shape *best);
/* All tiles x 180 sectors; `use_graph` replays the captured launch pair.
 * Asynchronous; surfaces accumulate as in tvs_run_sectors. */
tvs_status tvs_batch_run(tvs_batch *batch, uint32_t use_graph);
tvs_status tvs_batch_sync(tvs_batch *batch);
tvs_status tvs_batch_reset_surfaces(tvs_batch *batch);
tvs_status tvs_batch_read_surfaces(tvs_batch *batch, uint32_t tile, float *out,
                                   size_t n);
tvs_status tvs_batch_read_surfaces_f64(tvs_batch *batch, uint32_t tile, double *out,
                                       size_t n);
/* walk_ms = walk (or whole graph) time, reduce_ms = reduction time (0 when
 * the graph ran); point_steps over all tiles. */
tvs_status tvs_batch_get_timing(tvs_batch *batch, tvs_timing *out);
const char *tvs_batch_last_error(const tvs_batch *batch);

/* ---- host I/O around the walk ------------------------------------------ */

/* `.bt` header fields (input.rs:27-70, read at input.rs:90-113). */
typedef struct tvs_bt_header {
    uint32_t width;
    uint32_t height;
    uint16_t data_size;
    uint16_t is_float;
    uint16_t horizontal_units;
    uint16_t utm_zone;
    uint16_t datum;
    uint16_t projection_source;
    double left, right, bottom, top;
    float vertical_scale;
} tvs_bt_header;

/* `BinaryTerrain::read` header part (input.rs:86-113, 158-160): magic
 * "binterr1.3", fields, data-size check. */
tvs_status tvs_bt_read_header(const char *path, tvs_bt_header *out);

/* The samples after the 256-byte header (input.rs:133-171) as f32, widening
 * int16 exactly (main.rs:51-56).  `count` = width * height; the reference
 * then uses them as a width x width row-major grid (main.rs:49). */
tvs_status tvs_bt_read_f32(const char *path, float *out, size_t count);

/* The int16 samples without widening (int16 files only). */
tvs_status tvs_bt_read_i16(const char *path, int16_t *out, size_t count);

/* Write a `.bt` file (header + raw little-endian samples: int16, or f32 when
 * is_float), e.g. to hand a synthetic DEM to the CLI. */
tvs_status tvs_bt_write(const char *path, const tvs_bt_header *header,
                        const void *data);

/* `BinaryTerrain::scale` (input.rs:178-185): metres per point. */
double tvs_bt_scale(const tvs_bt_header *header);

/* `output::png::save` (output/png.rs:5-37): 8-bit greyscale PNG, min/max
 * normalised. */
tvs_status tvs_write_heatmap_png(const char *path, const float *data,
                                 uint32_t width, uint32_t height);

/* Band-data cache files `<state_dir>/total-viewsheds/<w>/{band_deltas,
 * distances}/<sector>.bin` (cache.rs:23-76): save a resident sector / load
 * one into the context (a file of another band length is rejected with
 * TVS_ERR_SIZE). */
tvs_status tvs_band_cache_save(tvs_ctx *ctx, const char *state_dir,
                               uint32_t sector);
tvs_status tvs_band_cache_load(tvs_ctx *ctx, const char *state_dir,
                               uint32_t sector);

/* Message for the last failing I/O call that had no context. */
const char *tvs_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TVS_H */
