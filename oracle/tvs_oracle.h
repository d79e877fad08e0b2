/*
 * tvs_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-precision-literal restatement of the reference
 * total-viewshed CPU path (tombh/total-viewsheds, `--compute cpu` as run by
 * its `#[cfg(test)]` build).  It exists to check the HIP engine: only the
 * tests, the smoke entry point and the benchmark's CPU-baseline leg may load
 * it.  The engine itself never links, loads or calls anything in oracle/.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference root:
 *   dem.rs, axes.rs, band_of_sight.rs, compute.rs, main.rs
 *       -> crates/total-viewsheds/src/
 *   kernel.rs -> crates/shader/src/
 *
 * What pins each function beyond retyping its formula (the reference's own
 * golden vectors, closed forms, brute force, constructed cases) is checked
 * by tests/test_oracle.py.
 */
#ifndef TVS_ORACLE_H
#define TVS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 0 is success, each non-zero value names the panic or `bail!`
 * site of the reference that the condition corresponds to. */
enum {
    TVSO_OK = 0,
    TVSO_ERR_LOS_TOO_LONG = 1,     /* dem.rs:56-62 bail!                       */
    TVSO_ERR_BAND_INDEX = 2,       /* band_of_sight.rs:103-168 index panics    */
    TVSO_ERR_RING_INDEX = 3,       /* kernel.rs:307 ring write out of bounds   */
    TVSO_ERR_DEM_INDEX = 4,        /* kernel.rs:164 elevation index panic      */
    TVSO_ERR_ALLOC = 5,
    TVSO_ERR_ARG = 6
};

/* dem.rs:11-40 (the scalar fields of `DEM`). */
typedef struct {
    uint32_t width;
    uint32_t size;
    float scale;
    uint32_t max_line_of_sight;
    uint32_t max_los_as_points;
    uint32_t computable_points_count;
    uint32_t tvs_width;
    uint32_t band_size;
} tvso_dem;

/* kernel.rs:38-58 `Constants` (the fields the kernel reads). */
typedef struct {
    uint32_t total_bands;
    uint32_t max_los_as_points;
    uint32_t dem_width;
    uint32_t tvs_width;
    float observer_height;
    uint32_t reserved_rings_per_band;
} tvso_constants;

/* dem.rs:44-84 `DEM::new` (+ count_computable_points 87-94,
 * is_point_computable 106-114, isqrt at 82). */
int tvso_dem_new(uint32_t width, float scale, uint32_t max_line_of_sight,
                 tvso_dem *out);

/* main.rs:36-38: default max line of sight `(w div 3) * scale` in f64. */
uint32_t tvso_default_max_line_of_sight(uint32_t width, double scale);

/* compute.rs:94-98 `ring_count_per_band`. */
uint32_t tvso_ring_count_per_band(float rings_per_km,
                                  uint32_t max_line_of_sight);

/* axes.rs:89-100: the shifted sector angle `angle + SECTOR_SHIFT` (f32). */
float tvso_shifted_angle(uint16_t sector);

/* axes.rs:154-176 `calculate_distances` (f64, row-major, y = -row). */
void tvso_calculate_distances(uint32_t width, float angle, double *out);

/* axes.rs:70-85 `order_by_distance`: stable ascending sort of ids by the
 * f32-rounded distance in IEEE total order (radsort 0.1.1 semantics). */
int tvso_order_by_distance(const double *distances, size_t n,
                           uint32_t *order);

/* axes.rs:103-130 `Axes::compute` for `Axes::new(width, sector)`:
 * fills axes.distances (f32, w*w), sight_ordered_map (w*w) and
 * sector_ordered (w*w). */
int tvso_axes_compute(uint32_t width, uint16_t sector, float *distances_f32,
                      uint32_t *sight_ordered_map, uint32_t *sector_ordered);

/* band_of_sight.rs:54-58 and 61-173 `compile_band_data`: writes
 * band_size-1 deltas and distances (zero tail included). */
int tvso_compile_band_data(const tvso_dem *dem, const float *axes_distances,
                           const uint32_t *sight_ordered_map,
                           const uint32_t *sector_ordered,
                           int32_t *band_deltas, float *band_distances);

/* compute.rs:118-161 with no cache: calculate_axes + compile_band_data for
 * one sector, allocating the axes internally. */
int tvso_band_data(const tvso_dem *dem, uint16_t sector, int32_t *band_deltas,
                   float *band_distances);

/* kernel.rs:71-256 `kernel` for one kernel_id.  `ring_data` may be NULL
 * (ring writes skipped, as on the SPIR-V build, kernel.rs:302-309).
 * Returns TVSO_OK or TVSO_ERR_RING_INDEX / TVSO_ERR_DEM_INDEX where the
 * reference would panic on an out-of-range index.  `*ring_spill` (if not
 * NULL) counts ring writes past the band's own `reserved_rings_per_band`
 * slots, which the reference lets run into the next band (kernel.rs:141). */
int tvso_kernel(uint32_t kernel_id, const tvso_constants *c,
                const float *elevations, size_t elevations_len,
                const float *distances, const int32_t *deltas,
                uint32_t n_deltas, float *cumulative_surfaces,
                uint32_t *ring_data, size_t ring_len, uint64_t *ring_spill);

/* compute.rs:27-84 (`Compute::new`, test-mode ring sizing at 46-47) and
 * compute.rs:101-115 + 181-240 (`run` with the CPU backend), for the sector
 * range [first_sector, first_sector + sector_count).
 *
 * `cumulative_surfaces` (computable_points_count floats) is accumulated
 * into, never cleared (compute.rs:102 clears once before sector 0: pass
 * zeros).  `ring_data`, if not NULL, receives sector_count consecutive
 * blocks of total_bands * R u32, each zeroed before its sector runs
 * (compute.rs:108).  With ring_data == NULL and n_threads > 1 the observers
 * are split over threads; each point's forward-then-backward order is kept,
 * so the result is bit-identical to the serial loop. */
int tvso_compute_run(const tvso_dem *dem, const float *elevations,
                     float observer_height, float rings_per_km,
                     uint16_t first_sector, uint16_t sector_count,
                     float *cumulative_surfaces, uint32_t *ring_data,
                     uint64_t *ring_spill, int n_threads);

/* The walk alone (compute.rs:228-240 for each sector) over precomputed band
 * data: `band_deltas`/`band_distances` hold sector_count blocks of
 * band_size-1 values.  Used to time the CPU baseline and as the checker of
 * the engine's walk when band data is supplied from outside. */
int tvso_walk_sectors(const tvso_dem *dem, const float *elevations,
                      float observer_height, uint32_t rings_per_band,
                      const int32_t *band_deltas, const float *band_distances,
                      uint16_t sector_count, float *cumulative_surfaces,
                      uint32_t *ring_data, int n_threads);

/* FP64 companion of the walk: the same band data and loop, with the height
 * difference, angle, max-scan and surface sum carried in double.  For every
 * point it accumulates the f64 surface over the sectors into tvs64, and it
 * counts the (sector, band, step) triples whose visibility differs from the
 * literal f32 walk (`*flips`) out of `*steps` evaluated. */
int tvso_walk_sectors_f64(const tvso_dem *dem, const float *elevations,
                          float observer_height, const int32_t *band_deltas,
                          const float *band_distances, uint16_t sector_count,
                          double *tvs64, uint64_t *flips, uint64_t *steps,
                          int n_threads);

#ifdef __cplusplus
}
#endif

#endif
