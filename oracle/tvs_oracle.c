/*
 * tvs_oracle.c -- TEST INFRASTRUCTURE ONLY (see tvs_oracle.h).
 *
 * Literal CPU restatement of the reference total-viewshed path.  Compile
 * with -ffp-contract=off and without -ffast-math: every f32/f64 operation
 * below must round exactly once, in the order the reference writes it.
 */
#include "tvs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* kernel.rs:13, 17, 21 */
#define TVSO_MAX_ANGLE (-2000.0f)
#define TVSO_TAN_ONE_RAD (0.0174533f)
#define TVSO_EARTH_RADIUS_SQUARED (12742000.0f)
/* axes.rs:35, 47 */
#define TVSO_SECTOR_STEPS 180u
#define TVSO_SECTOR_SHIFT (0.001f)

/* Rust `f as u32` (saturating; NaN -> 0). */
static uint32_t f32_to_u32_sat(float v)
{
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

static uint32_t f64_to_u32_sat(double v)
{
    if (!(v > 0.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}

/* `u32::isqrt` (dem.rs:82): floor of the real square root. */
static uint32_t isqrt_u32(uint32_t n)
{
    uint64_t r = (uint64_t)sqrt((double)n);
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return (uint32_t)r;
}

/* dem.rs:106-114 `is_point_computable`. */
static int is_point_computable(const tvso_dem *d, uint32_t dem_id)
{
    float los = (float)d->max_line_of_sight;
    float x = (float)(dem_id % d->width) * d->scale;
    float y = (float)(dem_id / d->width) * d->scale;
    float lower = los;
    float upper = ((float)(d->width - 1u) * d->scale) - los;
    return x >= lower && x <= upper && y >= lower && y <= upper;
}

/* dem.rs:44-84 */
int tvso_dem_new(uint32_t width, float scale, uint32_t max_line_of_sight,
                 tvso_dem *out)
{
    tvso_dem d;
    memset(&d, 0, sizeof d);
    d.width = width;
    d.size = width * width;                                   /* dem.rs:45 */
    d.scale = scale;
    d.max_line_of_sight = max_line_of_sight;
    d.max_los_as_points =
        f32_to_u32_sat((float)max_line_of_sight / scale);     /* dem.rs:53 */
    if (d.max_los_as_points > width / 2u)                     /* dem.rs:54-62 */
        return TVSO_ERR_LOS_TOO_LONG;
    d.band_size = d.max_los_as_points + 1u;                   /* dem.rs:78 */
    d.computable_points_count = 0;                            /* dem.rs:87-94 */
    for (uint32_t p = 0; p < d.size; ++p)
        if (is_point_computable(&d, p)) d.computable_points_count += 1u;
    d.tvs_width = isqrt_u32(d.computable_points_count);       /* dem.rs:82 */
    *out = d;
    return TVSO_OK;
}

/* main.rs:36-38 */
uint32_t tvso_default_max_line_of_sight(uint32_t width, double scale)
{
    return f64_to_u32_sat((double)(width / 3u) * scale);
}

/* compute.rs:94-98 */
uint32_t tvso_ring_count_per_band(float rings_per_km,
                                  uint32_t max_line_of_sight)
{
    float band_length_in_km = (float)max_line_of_sight / 1000.0f;
    return f32_to_u32_sat(band_length_in_km * rings_per_km);
}

/* axes.rs:89-100 */
float tvso_shifted_angle(uint16_t sector)
{
    return (float)sector + TVSO_SECTOR_SHIFT;
}

/* axes.rs:154-176.  Rust's f64::to_radians is `x * (PI / 180.0)`. */
void tvso_calculate_distances(uint32_t width, float angle, double *out)
{
    double angle_f64 = (double)angle;
    double sine_of_angle = sin(angle_f64 * (M_PI / 180.0));
    double cosine_of_angle = cos(angle_f64 * (M_PI / 180.0));
    size_t k = 0;
    for (int64_t y = 0; y >= -((int64_t)width - 1); --y) {
        for (uint32_t x = 0; x < width; ++x) {
            double left = (double)x * sine_of_angle;
            double right = (double)y * cosine_of_angle;
            out[k++] = left - right;
        }
    }
}

/* IEEE total order of an f32 as an unsigned key (-0 < +0). */
static uint32_t f32_total_order_key(float f)
{
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* axes.rs:70-85: the ids sorted by (f32)distance, ties kept in id order.
 * Written as an LSD radix sort over the total-order key (each byte pass is
 * a stable counting sort), which is the radsort 0.1.1 contract. */
int tvso_order_by_distance(const double *distances, size_t n, uint32_t *order)
{
    uint32_t *key = malloc(n * sizeof *key);
    uint32_t *key2 = malloc(n * sizeof *key2);
    uint32_t *id2 = malloc(n * sizeof *id2);
    if (!key || !key2 || !id2) {
        free(key); free(key2); free(id2);
        return TVSO_ERR_ALLOC;
    }
    for (size_t i = 0; i < n; ++i) {
        key[i] = f32_total_order_key((float)distances[i]);
        order[i] = (uint32_t)i;
    }
    uint32_t *ks = key, *is = order, *kd = key2, *id = id2;
    for (int pass = 0; pass < 4; ++pass) {
        size_t count[257];
        memset(count, 0, sizeof count);
        int shift = 8 * pass;
        for (size_t i = 0; i < n; ++i) count[((ks[i] >> shift) & 0xffu) + 1]++;
        for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
        for (size_t i = 0; i < n; ++i) {
            size_t dst = count[(ks[i] >> shift) & 0xffu]++;
            kd[dst] = ks[i];
            id[dst] = is[i];
        }
        uint32_t *t;
        t = ks; ks = kd; kd = t;
        t = is; is = id; id = t;
    }
    /* After 4 passes the result is back in `key`/`order`. */
    free(key); free(key2); free(id2);
    return TVSO_OK;
}

/* axes.rs:103-130 (+ convert_distances_to_f32 187-189). */
int tvso_axes_compute(uint32_t width, uint16_t sector, float *distances_f32,
                      uint32_t *sight_ordered_map, uint32_t *sector_ordered)
{
    size_t n = (size_t)width * width;
    float angle = tvso_shifted_angle(sector);
    double *dist = malloc(n * sizeof *dist);
    uint32_t *sight = malloc(n * sizeof *sight);
    if (!dist || !sight) {
        free(dist); free(sight);
        return TVSO_ERR_ALLOC;
    }
    tvso_calculate_distances(width, angle, dist);
    int st = tvso_order_by_distance(dist, n, sight);
    if (st == TVSO_OK) {
        for (size_t i = 0; i < n; ++i) distances_f32[i] = (float)dist[i];
        for (size_t i = 0; i < n; ++i) sight_ordered_map[sight[i]] = (uint32_t)i;
        tvso_calculate_distances(width, angle + 90.0f, dist);
        st = tvso_order_by_distance(dist, n, sector_ordered);
    }
    free(dist); free(sight);
    return st;
}

/* qsort has no context argument; one per thread, so band data of several
 * sectors can be compiled concurrently. */
static _Thread_local const uint32_t *g_sort_rank;

static int cmp_by_rank(const void *a, const void *b)
{
    uint32_t ra = g_sort_rank[*(const uint32_t *)a];
    uint32_t rb = g_sort_rank[*(const uint32_t *)b];
    return (ra > rb) - (ra < rb);
}

/* band_of_sight.rs:61-173 */
int tvso_compile_band_data(const tvso_dem *dem, const float *axes_distances,
                           const uint32_t *sight_ordered_map,
                           const uint32_t *sector_ordered,
                           int32_t *band_deltas, float *band_distances)
{
    size_t n_deltas = dem->band_size - 1u;                    /* 54-58 */
    for (size_t i = 0; i < n_deltas; ++i) {                   /* 67-68 */
        band_deltas[i] = 0;
        band_distances[i] = 0.0f;
    }
    size_t band_samples =
        (size_t)((double)(dem->band_size * 2u) * M_SQRT2);    /* 95 */
    size_t pov_id = dem->size / 2u;                           /* 99 */
    if (band_samples / 2 > pov_id) return TVSO_ERR_BAND_INDEX;
    size_t band_edge = pov_id - band_samples / 2;             /* 103 */
    if (band_edge + band_samples > dem->size) return TVSO_ERR_BAND_INDEX;

    uint32_t *dem_ids = malloc((band_samples + 1) * sizeof *dem_ids);
    uint32_t *distance_ids = malloc((band_samples + 1) * sizeof *distance_ids);
    uint32_t *order = malloc((band_samples + 1) * sizeof *order);
    if (!dem_ids || !distance_ids || !order) {
        free(dem_ids); free(distance_ids); free(order);
        return TVSO_ERR_ALLOC;
    }
    for (size_t i = 0; i < band_samples; ++i) {               /* 111-120 */
        uint32_t dem_id = sector_ordered[band_edge + i];
        dem_ids[i] = dem_id;
        distance_ids[i] = sight_ordered_map[dem_id];
    }
    /* 135-145: order the samples by their sight rank (ranks are distinct, so
     * the order is unique whatever the sort's stability). */
    for (size_t i = 0; i < band_samples; ++i) order[i] = (uint32_t)i;
    g_sort_rank = distance_ids;
    qsort(order, band_samples, sizeof *order, cmp_by_rank);

    size_t middle = band_samples / 2;                          /* 149 */
    size_t end_of_band = (size_t)dem->band_size * 2u;          /* 150 */
    float pov_distance = axes_distances[pov_id];               /* 157 */
    int st = TVSO_OK;
    for (size_t band_id = middle, i = 0; band_id < end_of_band; ++band_id, ++i) {
        if (i >= n_deltas || band_id + 1 >= band_samples) {    /* index panics */
            st = TVSO_ERR_BAND_INDEX;
            break;
        }
        int32_t current_dem_id = (int32_t)dem_ids[order[band_id]];
        int32_t next_dem_id = (int32_t)dem_ids[order[band_id + 1]];
        band_deltas[i] = current_dem_id - next_dem_id;          /* 164-165 */
        float next_distance = axes_distances[next_dem_id];     /* 167 */
        band_distances[i] = fabsf(pov_distance - next_distance); /* 168 */
    }
    free(dem_ids); free(distance_ids); free(order);
    return st;
}

int tvso_band_data(const tvso_dem *dem, uint16_t sector, int32_t *band_deltas,
                   float *band_distances)
{
    size_t n = (size_t)dem->width * dem->width;
    float *dist = malloc(n * sizeof *dist);
    uint32_t *sight_map = malloc(n * sizeof *sight_map);
    uint32_t *sector_ordered = malloc(n * sizeof *sector_ordered);
    int st = TVSO_ERR_ALLOC;
    if (dist && sight_map && sector_ordered) {
        st = tvso_axes_compute(dem->width, sector, dist, sight_map, sector_ordered);
        if (st == TVSO_OK)
            st = tvso_compile_band_data(dem, dist, sight_map, sector_ordered,
                                        band_deltas, band_distances);
    }
    free(dist); free(sight_map); free(sector_ordered);
    return st;
}

/* kernel.rs:302-309 `save_ring_data`. */
static int save_ring_data(uint32_t *ring, size_t ring_len, size_t index,
                          size_t value)
{
    if (!ring) return TVSO_OK;          /* SPIR-V build: writes compiled out */
    if (index >= ring_len) return TVSO_ERR_RING_INDEX;
    ring[index] = (uint32_t)value;
    return TVSO_OK;
}

/* kernel.rs:71-256 */
int tvso_kernel(uint32_t kernel_id, const tvso_constants *c,
                const float *elevations, size_t elevations_len,
                const float *distances, const int32_t *deltas,
                uint32_t n_deltas, float *cumulative_surfaces,
                uint32_t *ring_data, size_t ring_len, uint64_t *ring_spill)
{
    uint32_t tvs_id;
    float max_angle = TVSO_MAX_ANGLE;                       /* 102 */
    uint32_t half_total_bands = c->total_bands / 2u;        /* 103 */
    int is_currently_visible = 1;                           /* 104 */
    int closing = 0;                                        /* 105 */
    float band_surface = 0.0f;                              /* 108 */
    size_t ring_id = 1;                                     /* 114 */
    int forward;
    if (kernel_id < half_total_bands) {                     /* 117-123 */
        tvs_id = kernel_id;
        forward = 1;
    } else {
        tvs_id = kernel_id - half_total_bands;
        forward = 0;
    }
    uint32_t pov_x = tvs_id % c->tvs_width + c->max_los_as_points; /* 125 */
    uint32_t pov_y = tvs_id / c->tvs_width + c->max_los_as_points; /* 126 */
    size_t pov_id = (size_t)pov_y * c->dem_width + pov_x;          /* 131 */
    if (pov_id >= elevations_len) return TVSO_ERR_DEM_INDEX;
    float pov_elevation = elevations[pov_id] + c->observer_height; /* 134 */
    size_t ring_data_start =
        (size_t)kernel_id * c->reserved_rings_per_band;            /* 141 */
    int st;
    uint64_t spill = 0;
    if ((st = save_ring_data(ring_data, ring_len, ring_data_start + ring_id,
                             pov_id)) != TVSO_OK)                  /* 145 */
        return st;
    int is_previously_visible = 1;                                 /* 146 */
    size_t dem_id = pov_id;                                        /* 150 */

    for (uint32_t band_index = 0; band_index < n_deltas; ++band_index) {
        int32_t delta = deltas[band_index];
        /* 157-160 with delta_add / delta_subtract (259-292) */
        uint64_t absolute = (uint64_t)(delta < 0 ? -(int64_t)delta : delta);
        int add = forward ? (delta > 0) : !(delta > 0);
        if (add) {
            dem_id += absolute;
        } else {
            if (absolute > dem_id) return TVSO_ERR_DEM_INDEX;
            dem_id -= absolute;
        }
        if (dem_id >= elevations_len) return TVSO_ERR_DEM_INDEX;
        float elevation_delta = elevations[dem_id] - pov_elevation;  /* 164 */
        float distance = distances[band_index];                       /* 165 */
        float angle = (elevation_delta / distance) -
                      (distance / TVSO_EARTH_RADIUS_SQUARED);         /* 175 */
        is_currently_visible = angle > max_angle;                     /* 190 */
        int opening = is_currently_visible && !is_previously_visible; /* 194 */
        closing = is_previously_visible && !is_currently_visible;     /* 195 */
        if (is_currently_visible)                                     /* 197-217 */
            band_surface += distance * TVSO_TAN_ONE_RAD;
        if (opening) {                                                /* 220-223 */
            ring_id += 1;
            if (ring_id >= c->reserved_rings_per_band) ++spill;
            if ((st = save_ring_data(ring_data, ring_len,
                                     ring_data_start + ring_id, dem_id)) != TVSO_OK)
                return st;
        }
        if (closing) {                                                /* 226-229 */
            ring_id += 1;
            if (ring_id >= c->reserved_rings_per_band) ++spill;
            if ((st = save_ring_data(ring_data, ring_len,
                                     ring_data_start + ring_id, dem_id)) != TVSO_OK)
                return st;
        }
        is_previously_visible = is_currently_visible;                 /* 232 */
        max_angle = fmaxf(angle, max_angle);                          /* 233 */
    }
    if (is_currently_visible && !closing) {                           /* 237-240 */
        ring_id += 1;
        if (ring_id >= c->reserved_rings_per_band) ++spill;
        if ((st = save_ring_data(ring_data, ring_len,
                                 ring_data_start + ring_id, dem_id)) != TVSO_OK)
            return st;
    }
    if ((st = save_ring_data(ring_data, ring_len, ring_data_start,
                             ring_id)) != TVSO_OK)                    /* 243 */
        return st;
    cumulative_surfaces[tvs_id] += band_surface;                      /* 254 */
    if (ring_spill) *ring_spill += spill;
    return TVSO_OK;
}

static void make_constants(const tvso_dem *dem, float observer_height,
                           uint32_t rings_per_band, tvso_constants *c)
{
    c->total_bands = dem->computable_points_count * 2u;  /* compute.rs:34 */
    c->max_los_as_points = dem->max_los_as_points;
    c->dem_width = dem->width;
    c->tvs_width = dem->tvs_width;
    c->observer_height = observer_height;
    c->reserved_rings_per_band = rings_per_band;
}

/* compute.rs:228-240 for one sector. */
static int walk_one_sector(const tvso_constants *c, const float *elevations,
                           size_t elevations_len, const float *distances,
                           const int32_t *deltas, uint32_t n_deltas,
                           float *cum, uint32_t *ring, size_t ring_len,
                           uint64_t *spill, int n_threads)
{
    int st = TVSO_OK;
    uint32_t half = c->total_bands / 2u;
    if (ring || n_threads <= 1) {
        for (uint32_t k = 0; k < c->total_bands && st == TVSO_OK; ++k)
            st = tvso_kernel(k, c, elevations, elevations_len, distances, deltas,
                             n_deltas, cum, ring, ring_len, spill);
        return st;
    }
    /* Each point's two bands in the serial order (forward, then backward):
     * the only two kernels that touch cum[t]. */
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads) reduction(| : st)
    for (uint32_t t = 0; t < half; ++t) {
        int s1 = tvso_kernel(t, c, elevations, elevations_len, distances, deltas,
                             n_deltas, cum, NULL, 0, NULL);
        int s2 = tvso_kernel(t + half, c, elevations, elevations_len, distances,
                             deltas, n_deltas, cum, NULL, 0, NULL);
        st |= s1 | s2;
    }
    return st;
}

/* compute.rs:101-115, 181-203, 228-240 */
int tvso_compute_run(const tvso_dem *dem, const float *elevations,
                     float observer_height, float rings_per_km,
                     uint16_t first_sector, uint16_t sector_count,
                     float *cumulative_surfaces, uint32_t *ring_data,
                     uint64_t *ring_spill, int n_threads)
{
    tvso_constants c;
    uint32_t R = tvso_ring_count_per_band(rings_per_km, dem->max_line_of_sight);
    make_constants(dem, observer_height, R, &c);
    size_t ring_len = (size_t)c.total_bands * R;            /* compute.rs:47 */
    size_t n_deltas = dem->band_size - 1u;
    int32_t *deltas = malloc((n_deltas + 1) * sizeof *deltas);
    float *dists = malloc((n_deltas + 1) * sizeof *dists);
    if (!deltas || !dists) {
        free(deltas); free(dists);
        return TVSO_ERR_ALLOC;
    }
    int st = TVSO_OK;
    for (uint32_t s = first_sector;
         s < (uint32_t)first_sector + sector_count && st == TVSO_OK; ++s) {
        st = tvso_band_data(dem, (uint16_t)s, deltas, dists);
        if (st != TVSO_OK) break;
        uint32_t *ring = NULL;
        if (ring_data) {
            ring = ring_data + (size_t)(s - first_sector) * ring_len;
            memset(ring, 0, ring_len * sizeof *ring);        /* compute.rs:108 */
        }
        st = walk_one_sector(&c, elevations, dem->size, dists, deltas,
                             (uint32_t)n_deltas, cumulative_surfaces, ring,
                             ring_len, ring_spill, n_threads);
    }
    free(deltas); free(dists);
    return st;
}

int tvso_walk_sectors(const tvso_dem *dem, const float *elevations,
                      float observer_height, uint32_t rings_per_band,
                      const int32_t *band_deltas, const float *band_distances,
                      uint16_t sector_count, float *cumulative_surfaces,
                      uint32_t *ring_data, int n_threads)
{
    tvso_constants c;
    make_constants(dem, observer_height, rings_per_band, &c);
    size_t ring_len = (size_t)c.total_bands * rings_per_band;
    size_t n_deltas = dem->band_size - 1u;
    int st = TVSO_OK;
    for (uint32_t s = 0; s < sector_count && st == TVSO_OK; ++s) {
        uint32_t *ring = NULL;
        if (ring_data) {
            ring = ring_data + (size_t)s * ring_len;
            memset(ring, 0, ring_len * sizeof *ring);
        }
        st = walk_one_sector(&c, elevations, dem->size,
                             band_distances + (size_t)s * n_deltas,
                             band_deltas + (size_t)s * n_deltas,
                             (uint32_t)n_deltas, cumulative_surfaces, ring,
                             ring_len, NULL, n_threads);
    }
    return st;
}

/* One band in f32 (literal, kernel.rs:153-234) and f64 side by side.
 * Returns the f64 band surface; adds visibility disagreements to *flips. */
static double band_f64(const tvso_constants *c, const float *elev,
                       size_t pov_id, int forward, const float *dists,
                       const int32_t *deltas, uint32_t n, uint64_t *flips,
                       int *st)
{
    float h0 = elev[pov_id] + c->observer_height;
    double h0d = (double)elev[pov_id] + (double)c->observer_height;
    float mx = TVSO_MAX_ANGLE;
    double mxd = (double)TVSO_MAX_ANGLE;
    double surf = 0.0;
    size_t id = pov_id;
    uint64_t f = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t delta = deltas[i];
        int64_t step = forward ? (int64_t)delta : -(int64_t)delta;
        int64_t nid = (int64_t)id + step;
        if (nid < 0 || (uint64_t)nid >= (uint64_t)c->dem_width * c->dem_width) {
            *st = TVSO_ERR_DEM_INDEX;
            return 0.0;
        }
        id = (size_t)nid;
        float d = dists[i];
        float ang = ((elev[id] - h0) / d) - (d / TVSO_EARTH_RADIUS_SQUARED);
        double dd = (double)d;
        double angd = (((double)elev[id] - h0d) / dd) -
                      (dd / (double)TVSO_EARTH_RADIUS_SQUARED);
        int vis = ang > mx;
#ifdef TVSO_PLANTED_F32_VISIBILITY
        /* A planted mistake, compiled only by tests/test_oracle_sanitized.py
         * to show tests/test_oracle.py catches it: the f64 walk takes its
         * visibility from the f32 angle. */
        int visd = vis;
#else
        int visd = angd > mxd;
#endif
        f += (uint64_t)(vis != visd);
        if (visd) surf += dd * 0.0174533;
        mx = fmaxf(ang, mx);
        mxd = fmax(angd, mxd);
    }
    *flips += f;
    return surf;
}

int tvso_walk_sectors_f64(const tvso_dem *dem, const float *elevations,
                          float observer_height, const int32_t *band_deltas,
                          const float *band_distances, uint16_t sector_count,
                          double *tvs64, uint64_t *flips, uint64_t *steps,
                          int n_threads)
{
    tvso_constants c;
    make_constants(dem, observer_height, 0, &c);
    uint32_t n = dem->band_size - 1u;
    uint32_t N = dem->computable_points_count;
    uint64_t total_flips = 0;
    int st_all = TVSO_OK;
    if (n_threads < 1) n_threads = 1;
    for (uint32_t s = 0; s < sector_count; ++s) {
        const int32_t *dl = band_deltas + (size_t)s * n;
        const float *ds = band_distances + (size_t)s * n;
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads) reduction(+ : total_flips) reduction(| : st_all)
        for (uint32_t t = 0; t < N; ++t) {
            int st = TVSO_OK;
            uint32_t px = t % c.tvs_width + c.max_los_as_points;
            uint32_t py = t / c.tvs_width + c.max_los_as_points;
            size_t pov = (size_t)py * c.dem_width + px;
            uint64_t f = 0;
            double fwd = band_f64(&c, elevations, pov, 1, ds, dl, n, &f, &st);
            double bwd = band_f64(&c, elevations, pov, 0, ds, dl, n, &f, &st);
            tvs64[t] += fwd;
            tvs64[t] += bwd;
            total_flips += f;
            st_all |= st;
        }
    }
    if (flips) *flips = total_flips;
    if (steps) *steps = (uint64_t)sector_count * 2u * N * n;
    return st_all;
}
