// bands.hip -- per-sector band data on the device: the rotated-coordinate
// distances of every DEM point (axes.rs:154-176, f64), the two stable
// orderings of `Axes::compute` (axes.rs:103-130) as LSD radix sorts of
// 32-bit total-order keys, the band template compile
// (band_of_sight.rs:61-173) and the walk's per-step constant tables.
#include "engine.hpp"
#include "fastdiv.hpp"

#pragma clang fp contract(off)

namespace tvs {
namespace {

constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortItems = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kCompileThreads = 1024;
constexpr uint32_t kMaxBandSamples = 8192;

// f32 -> u32 whose unsigned order is the IEEE total order (-0 < +0): the
// float key order of radsort 0.1.1 (axes.rs:82).
__device__ __forceinline__ uint32_t total_order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// axes.rs:300-313: row r has y = -r (an i64 cast to f64), column x;
// distance = fl64(x*sin) - fl64(y*cos).
__device__ __forceinline__ double rotated_distance(uint32_t id, uint32_t w,
                                                   double s, double c) {
    const uint32_t r = id / w;
    const uint32_t x = id - r * w;
    const double y = static_cast<double>(-static_cast<int64_t>(r));
    const double left = static_cast<double>(x) * s;
    const double right = y * c;
    return left - right;
}

__global__ void keys_kernel(uint32_t w, double s, double c,
                            uint32_t *__restrict__ key,
                            uint32_t *__restrict__ id,
                            float *__restrict__ dist_f32) {
    const uint32_t n = w * w;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        const float d = static_cast<float>(rotated_distance(i, w, s, c));
        key[i] = total_order_key(d);
        id[i] = i;
        if (dist_f32) dist_f32[i] = d;
    }
}

// LSD radix sort, 8-bit digits.  Pass = histogram per tile -> one exclusive
// scan over the digit-major [256][tiles] table -> stable scatter.
__global__ __launch_bounds__(kSortThreads) void sort_hist_kernel(
    const uint32_t *__restrict__ key, uint32_t n, uint32_t shift,
    uint32_t *__restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t h[4][256];
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += kSortThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t i = base + j * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[wave][(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    const uint32_t d = threadIdx.x;
    hist[d * tiles + blockIdx.x] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

__global__ __launch_bounds__(kScanThreads) void sort_scan_kernel(
    uint32_t *__restrict__ data, uint32_t m) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t chunk = (m + kScanThreads - 1) / kScanThreads;
    const uint32_t begin = threadIdx.x * chunk;
    const uint32_t end = begin + chunk < m ? begin + chunk : m;
    uint32_t local = 0;
    for (uint32_t i = begin; i < end; ++i) local += data[i];
    part[threadIdx.x] = local;
    __syncthreads();
    for (uint32_t o = 1; o < kScanThreads; o <<= 1) {
        const uint32_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t i = begin; i < end; ++i) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    const uint32_t *__restrict__ key_in, const uint32_t *__restrict__ id_in,
    uint32_t *__restrict__ key_out, uint32_t *__restrict__ id_out, uint32_t n,
    uint32_t shift, const uint32_t *__restrict__ offsets, uint32_t tiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wave_count[4][256];
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t lane = tid & 63u;
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    run[tid] = offsets[tid * tiles + blockIdx.x];
    for (uint32_t wv = 0; wv < 4; ++wv) wave_count[wv][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t i = base + j * kSortThreads + tid;
        const bool valid = i < n;
        const uint32_t k = valid ? key_in[i] : 0u;
        const uint32_t v = valid ? id_in[i] : 0u;
        const uint32_t digit = (k >> shift) & 255u;
        // Lanes of this wave holding the same digit, from 8 ballots.
        uint64_t peers = __ballot(valid);
        for (uint32_t bit = 0; bit < 8; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const uint64_t bal = __ballot(set);
            peers &= set ? bal : ~bal;
        }
        const uint32_t rank = __popcll(peers & lanes_below);
        if (valid && rank == 0) wave_count[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[digit] + rank;
            for (uint32_t wv = 0; wv < wave; ++wv) pos += wave_count[wv][digit];
            key_out[pos] = k;
            id_out[pos] = v;
        }
        __syncthreads();
        run[tid] += wave_count[0][tid] + wave_count[1][tid] + wave_count[2][tid] +
                    wave_count[3][tid];
        for (uint32_t wv = 0; wv < 4; ++wv) wave_count[wv][tid] = 0;
        __syncthreads();
    }
}

__global__ void inverse_map_kernel(const uint32_t *__restrict__ order, uint32_t n,
                                   uint32_t *__restrict__ rank) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        rank[order[i]] = i;
}

// band_of_sight.rs:95-169 for one sector.  The window of `band_samples`
// points around the middle of the sector order is put in sight order
// (ranks are distinct, so a bitonic sort of (rank, sample) gives the
// reference's order); deltas and |D(P) - D(next)| fill the first
// 2B - bs/2 steps, the rest stay 0 / +0.
__global__ __launch_bounds__(kCompileThreads) void band_compile_kernel(
    const uint32_t *__restrict__ sector_ordered,
    const uint32_t *__restrict__ sight_map, uint32_t w, uint32_t band_size,
    uint32_t band_samples, uint32_t nb, double s_sight, double c_sight,
    int32_t *__restrict__ deltas, float *__restrict__ dists) {
    extern __shared__ uint64_t sample[];
    const uint32_t tid = threadIdx.x;
    const uint32_t pov_id = (w * w) / 2u;                    // 99
    const uint32_t band_edge = pov_id - band_samples / 2u;   // 103
    for (uint32_t j = tid; j < nb; j += kCompileThreads) {   // 111-120
        uint64_t v = ~0ull;
        if (j < band_samples) {
            const uint32_t dem_id = sector_ordered[band_edge + j];
            v = (static_cast<uint64_t>(sight_map[dem_id]) << 32) | j;
        }
        sample[j] = v;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= nb; k <<= 1) {                 // 135-145
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < nb; i += kCompileThreads) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = sample[i], b = sample[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        sample[i] = b;
                        sample[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint32_t middle = band_samples / 2u;               // 149
    const uint32_t end_of_band = band_size * 2u;             // 150
    const uint32_t n_steps = band_size - 1u;
    const float pov_distance =
        static_cast<float>(rotated_distance(pov_id, w, s_sight, c_sight));  // 157
    for (uint32_t i = tid; i < n_steps; i += kCompileThreads) {
        int32_t delta = 0;
        float distance = 0.0f;
        const uint32_t band_id = middle + i;
        if (band_id < end_of_band) {                         // 158-169
            const uint32_t cur = sector_ordered[band_edge + static_cast<uint32_t>(sample[band_id])];
            const uint32_t next = sector_ordered[band_edge + static_cast<uint32_t>(sample[band_id + 1])];
            delta = static_cast<int32_t>(cur) - static_cast<int32_t>(next);
            const float next_distance =
                static_cast<float>(rotated_distance(next, w, s_sight, c_sight));
            distance = fabsf(pov_distance - next_distance);
        }
        deltas[i] = delta;
        dists[i] = distance;
    }
}

// Per sector slot: cumulative offsets (block scan), 1/d, curvature term,
// surface weight, and the SectorMeta ranges.
__global__ __launch_bounds__(kScanThreads) void step_table_kernel(
    StepTables tab, uint32_t first_slot, uint32_t n_steps) {
    __shared__ int32_t part[kScanThreads];
    __shared__ int32_t red_min[kScanThreads / 64], red_max[kScanThreads / 64];
    __shared__ uint32_t red_fast[kScanThreads / 64], red_tail[kScanThreads / 64];
    const uint32_t slot = first_slot + blockIdx.x;
    const size_t tb = static_cast<size_t>(slot) * tab.stride;
    const int32_t *delta = tab.delta + tb;
    const float *dist = tab.dist + tb;
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (n_steps + kScanThreads - 1) / kScanThreads;
    const uint32_t begin = tid * chunk;
    const uint32_t end = begin + chunk < n_steps ? begin + chunk : n_steps;
    int32_t local = 0;
    for (uint32_t i = begin; i < end; ++i) local += delta[i];
    part[tid] = local;
    __syncthreads();
    for (uint32_t o = 1; o < kScanThreads; o <<= 1) {
        const int32_t v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int32_t run = tid ? part[tid - 1] : 0;
    int32_t mn = 0x7fffffff, mx = -0x7fffffff - 1;
    uint32_t first_irregular = n_steps;
    uint32_t last_nonzero = 0;  // one past the last step outside the zero tail
    for (uint32_t i = begin; i < end; ++i) {
        run += delta[i];
        const float d = dist[i];
        tab.off[tb + i] = run;
        const Reciprocal y = split_reciprocal(d);
        tab.rcp[tb + i] = y.hi;
        tab.rcp_lo[tb + i] = y.lo;
        tab.curv[tb + i] = d / kEarthRadiusSquared;
        tab.wt[tb + i] = d * kTanOneRad;
        mn = run < mn ? run : mn;
        mx = run > mx ? run : mx;
        if (!regular_distance(d) && i < first_irregular) first_irregular = i;
        if (delta[i] != 0 || __float_as_uint(d) != 0u) last_nonzero = i + 1;
    }
    // Wave then block reductions.
    for (uint32_t o = 32; o > 0; o >>= 1) {
        mn = min(mn, __shfl_xor(mn, o));
        mx = max(mx, __shfl_xor(mx, o));
        first_irregular = min(first_irregular, static_cast<uint32_t>(__shfl_xor(static_cast<int>(first_irregular), o)));
        last_nonzero = max(last_nonzero, static_cast<uint32_t>(__shfl_xor(static_cast<int>(last_nonzero), o)));
    }
    if ((tid & 63u) == 0) {
        red_min[tid >> 6] = mn;
        red_max[tid >> 6] = mx;
        red_fast[tid >> 6] = first_irregular;
        red_tail[tid >> 6] = last_nonzero;
    }
    __syncthreads();
    if (tid == 0) {
        SectorMeta m{};
        m.min_off = 0;   // the PoV itself is on every band
        m.max_off = 0;
        m.n_fast = n_steps;
        m.tail_start = 0;
        for (uint32_t wv = 0; wv < kScanThreads / 64; ++wv) {
            m.min_off = min(m.min_off, red_min[wv]);
            m.max_off = max(m.max_off, red_max[wv]);
            m.n_fast = min(m.n_fast, red_fast[wv]);
            m.tail_start = max(m.tail_start, red_tail[wv]);
        }
        if (m.n_fast > m.tail_start) m.n_fast = m.tail_start;
        m.status = 0;
        tab.meta[slot] = m;
    }
}

uint32_t grid_for(size_t n, uint32_t block) {
    size_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return static_cast<uint32_t>(g ? g : 1);
}

hipError_t radix_sort(tvs_ctx *c, size_t n) {
    SortScratch &s = c->sort;
    const uint32_t tiles = static_cast<uint32_t>((n + kSortTile - 1) / kSortTile);
    uint32_t *ks = s.key_a, *is = s.id_a, *kd = s.key_b, *id = s.id_b;
    for (uint32_t pass = 0; pass < 4; ++pass) {
        const uint32_t shift = 8u * pass;
        hipLaunchKernelGGL(sort_hist_kernel, dim3(tiles), dim3(kSortThreads), 0, c->stream,
                           ks, static_cast<uint32_t>(n), shift, s.hist, tiles);
        hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(kScanThreads), 0, c->stream,
                           s.hist, 256u * tiles);
        hipLaunchKernelGGL(sort_scatter_kernel, dim3(tiles), dim3(kSortThreads), 0,
                           c->stream, ks, is, kd, id, static_cast<uint32_t>(n), shift,
                           s.hist, tiles);
        std::swap(ks, kd);
        std::swap(is, id);
    }
    return hipGetLastError();  // 4 passes: the result is back in key_a / id_a
}

}  // namespace

// Device Axes::compute for one sector.  Afterwards: sort.sight_map holds the
// inverse sight order, sort.id_a the sector order, sort.dist_f32 (if asked)
// axes.distances.
hipError_t launch_axes(tvs_ctx *c, uint32_t sector, bool want_distances) {
    const uint32_t w = c->geo.dem_width;
    const size_t n = static_cast<size_t>(w) * w;
    SortScratch &s = c->sort;
    double sn, cs;
    sector_sin_cos(sector, false, &sn, &cs);
    hipLaunchKernelGGL(keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, w,
                       sn, cs, s.key_a, s.id_a, want_distances ? s.dist_f32 : nullptr);
    hipError_t e = radix_sort(c, n);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(inverse_map_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                       c->stream, s.id_a, static_cast<uint32_t>(n), s.sight_map);
    sector_sin_cos(sector, true, &sn, &cs);
    hipLaunchKernelGGL(keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, w,
                       sn, cs, s.key_a, s.id_a, nullptr);
    e = radix_sort(c, n);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_band_compile(tvs_ctx *c, uint32_t sector) {
    const tvs_geometry &g = c->geo;
    uint32_t nb = 1;
    while (nb < g.band_samples) nb <<= 1;
    if (nb > kMaxBandSamples) return hipErrorInvalidValue;
    double sn, cs;
    sector_sin_cos(sector, false, &sn, &cs);
    const size_t tb = static_cast<size_t>(sector) * c->tab.stride;
    hipLaunchKernelGGL(band_compile_kernel, dim3(1), dim3(kCompileThreads),
                       nb * sizeof(uint64_t), c->stream, c->sort.id_a, c->sort.sight_map,
                       g.dem_width, g.band_size, g.band_samples, nb, sn, cs,
                       c->tab.delta + tb, c->tab.dist + tb);
    return hipGetLastError();
}

hipError_t launch_step_tables(tvs_ctx *c, uint32_t first_slot, uint32_t count) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(step_table_kernel, dim3(count), dim3(kScanThreads), 0, c->stream,
                       c->tab, first_slot, c->geo.band_deltas_size);
    return hipGetLastError();
}

uint32_t max_band_samples() { return kMaxBandSamples; }

}  // namespace tvs
