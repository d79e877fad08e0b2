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

// LSD radix sort, 8-bit digits.  Pass = histogram per tile -> per-digit
// exclusive scans over the digit-major [256][tiles] table (+ 256 totals) ->
// stable scatter.
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

// One block per digit: exclusive scan of that digit's per-tile counts (in
// place) and the digit's total.
__global__ __launch_bounds__(kScanThreads) void sort_rowscan_kernel(
    uint32_t *__restrict__ hist, uint32_t tiles, uint32_t *__restrict__ totals) {
    __shared__ uint32_t part[kScanThreads];
    uint32_t *row = hist + static_cast<size_t>(blockIdx.x) * tiles;
    const uint32_t chunk = (tiles + kScanThreads - 1) / kScanThreads;
    const uint32_t begin = threadIdx.x * chunk;
    const uint32_t end = begin + chunk < tiles ? begin + chunk : tiles;
    uint32_t local = 0;
    for (uint32_t i = begin; i < end; ++i) local += row[i];
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
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == kScanThreads - 1) totals[blockIdx.x] = part[kScanThreads - 1];
}

__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    const uint32_t *__restrict__ key_in, const uint32_t *__restrict__ id_in,
    uint32_t *__restrict__ key_out, uint32_t *__restrict__ id_out, uint32_t n,
    uint32_t shift, const uint32_t *__restrict__ offsets, uint32_t tiles,
    const uint32_t *__restrict__ totals) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wave_count[4][256];
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t lane = tid & 63u;
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    // Start of digit `tid` in the output: exclusive scan of the digit totals.
    run[tid] = totals[tid];
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
        const uint32_t v = tid >= o ? run[tid - o] : 0u;
        __syncthreads();
        run[tid] += v;
        __syncthreads();
    }
    const uint32_t digit_base = run[tid] - totals[tid];
    __syncthreads();
    run[tid] = digit_base + offsets[tid * tiles + blockIdx.x];
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
    StepTables tab, uint32_t first_slot, uint32_t n_steps, uint32_t copy_offset,
    uint32_t parity) {
    __shared__ int32_t part[kScanThreads];
    __shared__ int32_t red_min[kScanThreads / 64], red_max[kScanThreads / 64];
    __shared__ uint32_t red_fast[kScanThreads / 64], red_tail[kScanThreads / 64];
    __shared__ uint32_t red_reg[kScanThreads / 64], red_fast2[kScanThreads / 64];
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
    uint32_t first_regular = n_steps;
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
        if (regular_distance(d) && i < first_regular) first_regular = i;
        if (delta[i] != 0 || __float_as_uint(d) != 0u) last_nonzero = i + 1;
    }
    // Wave then block reductions.
    for (uint32_t o = 32; o > 0; o >>= 1) {
        mn = min(mn, __shfl_xor(mn, o));
        mx = max(mx, __shfl_xor(mx, o));
        first_irregular = min(first_irregular, static_cast<uint32_t>(__shfl_xor(static_cast<int>(first_irregular), o)));
        first_regular = min(first_regular, static_cast<uint32_t>(__shfl_xor(static_cast<int>(first_regular), o)));
        last_nonzero = max(last_nonzero, static_cast<uint32_t>(__shfl_xor(static_cast<int>(last_nonzero), o)));
    }
    if ((tid & 63u) == 0) {
        red_min[tid >> 6] = mn;
        red_max[tid >> 6] = mx;
        red_fast[tid >> 6] = first_irregular;
        red_reg[tid >> 6] = first_regular;
        red_tail[tid >> 6] = last_nonzero;
    }
    __syncthreads();
    // Block-wide results (every thread needs the bias).
    int32_t lo = 0, hi = 0;   // the PoV itself is on every band
    uint32_t fast = n_steps, tail = 0, reg = n_steps;
    for (uint32_t wv = 0; wv < kScanThreads / 64; ++wv) {
        lo = min(lo, red_min[wv]);
        hi = max(hi, red_max[wv]);
        fast = min(fast, red_fast[wv]);
        reg = min(reg, red_reg[wv]);
        tail = max(tail, red_tail[wv]);
    }
    // Leading irregular steps (a first band point at distance +0, say) are
    // walked one at a time before the fast range: head = the first regular
    // step when the band opens with irregular ones, and the fast range ends
    // at the first irregular step after it.
    const uint32_t head = fast < reg ? reg : 0u;
    if (head != 0u) {
        uint32_t irr = n_steps;
        for (uint32_t i = begin > head ? begin : head; i < end; ++i)
            if (!regular_distance(dist[i])) {
                irr = i;
                break;
            }
        for (uint32_t o = 32; o > 0; o >>= 1)
            irr = min(irr, static_cast<uint32_t>(__shfl_xor(static_cast<int>(irr), o)));
        if ((tid & 63u) == 0) red_fast2[tid >> 6] = irr;
    }
    __syncthreads();
    if (head != 0u) {
        fast = n_steps;
        for (uint32_t wv = 0; wv < kScanThreads / 64; ++wv) fast = min(fast, red_fast2[wv]);
    }
    // Even, so that the buffer base 2 * bias bytes below the DEM and every
    // position below are dword-aligned wherever the element they reach is.
    int32_t bias = max(hi, -lo);
    bias += bias & 1;
    // Byte offsets of int16 elements: 2 * (bias + parity +/- off), plus
    // 2 * copy_offset (copy B; 0 when the DEM has none) where a two-point
    // pair's element pov +/- off is odd (pov has the pairs' parity; off and
    // -off have one parity).  The walks' vector offset is that of element
    // pov - parity, so the sum reaches pov +/- off, and both offsets of a
    // two-point load are dword-aligned.  The host keeps them below 2^32.
    for (uint32_t i = begin; i < end; ++i) {
        const int32_t o = tab.off[tb + i];
        const uint32_t b = ((static_cast<uint32_t>(o) + parity) & 1u) ? 2u * copy_offset : 0u;
        tab.pos_f[tb + i] = 2u * (static_cast<uint32_t>(bias + o) + parity) + b;
        tab.pos_b[tb + i] = 2u * (static_cast<uint32_t>(bias - o) + parity) + b;
    }
    if (tid == 0) {
        SectorMeta m{};
        m.min_off = lo;
        m.max_off = hi;
        m.n_fast = fast > tail ? tail : fast;
        m.head = head > m.n_fast ? m.n_fast : head;
        m.tail_start = tail;
        m.status = 0;
        m.bias = static_cast<uint32_t>(bias);
        tab.meta[slot] = m;
    }
}


// ---- band data without the full sorts --------------------------------------
//
// compile_band_data (band_of_sight.rs:95-169) reads two things from the axes:
// the `band_samples` points at ranks [P - bs/2, P - bs/2 + bs) of the sector
// order, and the relative sight order of those samples.  Both orders sort by
// (f32 key in IEEE total order, id), so:
//  * within a DEM row the sector key is monotone in x (x*sin - y*cos with y
//    fixed; every rounding step is monotone), which makes
//    count(key <= K) a sum of per-row binary searches, and the keys at the
//    window's first and last rank follow by bisection over the 32-bit key;
//  * the points with key between those two keys are few (the window plus
//    ties), sorted by (key, id) in LDS; their global ranks start at
//    count(key < K_lo), which cuts the window out exactly;
//  * the samples' sight order is their (sight key, id) order.
// One block per sector; all sectors of a call in one launch.

constexpr uint32_t kSelectThreads = 1024;
constexpr uint32_t kSelectCapacity = 8192;   // candidates held in LDS

struct SectorAxes {
    double s_sight, c_sight;     // sight axis (a = sector + 0.001)
    double s_sector, c_sector;   // sector axis (a + 90)
};

__device__ __forceinline__ uint32_t sector_key(uint32_t r, uint32_t x, double s, double c) {
    const double y = static_cast<double>(-static_cast<int64_t>(r));
    const double left = static_cast<double>(x) * s;
    const double right = y * c;
    return total_order_key(static_cast<float>(left - right));
}

// Number of x in [0, w) of row r whose key is <= K (rows are non-decreasing
// in x when s >= 0, non-increasing otherwise).
__device__ __forceinline__ uint32_t row_count_le(uint32_t r, uint32_t w, double s, double c,
                                                 bool increasing, uint32_t K) {
    uint32_t lo = 0, hi = w;
    if (increasing) {
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sector_key(r, mid, s, c) <= K) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sector_key(r, mid, s, c) <= K) hi = mid; else lo = mid + 1;
    }
    return w - lo;
}

__device__ uint32_t block_sum(uint32_t v, uint32_t *scratch) {
    for (uint32_t o = 32; o > 0; o >>= 1) v += __shfl_xor(static_cast<int>(v), o);
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) scratch[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t i = 0; i < kSelectThreads / 64; ++i) t += scratch[i];
    return t;
}

__device__ void bitonic_sort(uint64_t *a, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t x = a[i], y = a[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kSelectThreads) void band_select_kernel(
    const SectorAxes *__restrict__ axes, uint32_t first_sector, uint32_t w,
    uint32_t band_size, uint32_t band_samples, int32_t *__restrict__ deltas_base,
    float *__restrict__ dists_base, uint32_t stride, uint32_t *__restrict__ status) {
    __shared__ uint64_t cand[kSelectCapacity];
    __shared__ uint32_t red[kSelectThreads / 64];
    __shared__ uint32_t row_base[kSelectThreads + 1];
    const uint32_t sector = first_sector + blockIdx.x;
    const SectorAxes ax = axes[sector];
    const double s = ax.s_sector, c = ax.c_sector;
    const bool inc = s >= 0.0;
    const uint32_t tid = threadIdx.x;
    const uint32_t n = w * w;
    const uint32_t pov_id = n / 2u;
    const uint32_t edge = pov_id - band_samples / 2u;
    const uint32_t last = edge + band_samples - 1u;

    // Keys at ranks `edge` and `last`: smallest K with count(key <= K) > rank.
    uint32_t lo1 = 0, hi1 = 0xffffffffu, lo2 = 0, hi2 = 0xffffffffu;
    while (lo1 < hi1 || lo2 < hi2) {
        const uint32_t m1 = lo1 + ((hi1 - lo1) >> 1);
        const uint32_t m2 = lo2 + ((hi2 - lo2) >> 1);
        uint32_t c1 = 0, c2 = 0;
        for (uint32_t r = tid; r < w; r += kSelectThreads) {
            if (lo1 < hi1) c1 += row_count_le(r, w, s, c, inc, m1);
            if (lo2 < hi2) c2 += row_count_le(r, w, s, c, inc, m2);
        }
        c1 = block_sum(c1, red);
        c2 = block_sum(c2, red);
        if (lo1 < hi1) { if (c1 > edge) hi1 = m1; else lo1 = m1 + 1; }
        if (lo2 < hi2) { if (c2 > last) hi2 = m2; else lo2 = m2 + 1; }
    }
    const uint32_t k_lo = lo1, k_hi = lo2;
    // Ranks start at count(key < k_lo).
    uint32_t below = 0;
    if (k_lo > 0)
        for (uint32_t r = tid; r < w; r += kSelectThreads)
            below += row_count_le(r, w, s, c, inc, k_lo - 1u);
    below = block_sum(below, red);

    // Candidates: per row the x range with k_lo <= key <= k_hi.  Rows are
    // split in contiguous chunks per thread so ids come out row-major.
    const uint32_t rows_per = (w + kSelectThreads - 1) / kSelectThreads;
    const uint32_t r0 = tid * rows_per < w ? tid * rows_per : w;
    const uint32_t r1 = r0 + rows_per < w ? r0 + rows_per : w;
    uint32_t mine = 0;
    for (uint32_t r = r0; r < r1; ++r) {
        const uint32_t le_hi = row_count_le(r, w, s, c, inc, k_hi);
        const uint32_t lt_lo = k_lo ? row_count_le(r, w, s, c, inc, k_lo - 1u) : 0u;
        mine += le_hi - lt_lo;
    }
    row_base[tid + 1] = mine;
    if (tid == 0) row_base[0] = 0;
    __syncthreads();
    if (tid == 0)
        for (uint32_t i = 1; i <= kSelectThreads; ++i) row_base[i] += row_base[i - 1];
    __syncthreads();
    const uint32_t total = row_base[kSelectThreads];
    uint32_t nb = 1;
    while (nb < total) nb <<= 1;
    if (nb > kSelectCapacity) {
        if (tid == 0) status[sector] = 1u;   // host takes the full-sort path
        return;
    }
    uint32_t pos = row_base[tid];
    for (uint32_t r = r0; r < r1; ++r) {
        const uint32_t le_hi = row_count_le(r, w, s, c, inc, k_hi);
        const uint32_t lt_lo = k_lo ? row_count_le(r, w, s, c, inc, k_lo - 1u) : 0u;
        // Increasing rows: x in [lt_lo, le_hi); decreasing: [w-le_hi, w-lt_lo).
        const uint32_t xa = inc ? lt_lo : w - le_hi;
        const uint32_t xb = inc ? le_hi : w - lt_lo;
        for (uint32_t x = xa; x < xb; ++x)
            cand[pos++] = (static_cast<uint64_t>(sector_key(r, x, s, c)) << 32) | (r * w + x);
    }
    for (uint32_t i = total + tid; i < nb; i += kSelectThreads) cand[i] = ~0ull;
    __syncthreads();
    bitonic_sort(cand, nb);

    // The window, re-keyed by sight order: (sight key, id).
    const uint32_t off = edge - below;
    uint32_t nb2 = 1;
    while (nb2 < band_samples) nb2 <<= 1;
    constexpr uint32_t kPer = kSelectCapacity / kSelectThreads;
    uint64_t held[kPer];
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t j = tid + q * kSelectThreads;
        held[q] = j < band_samples ? cand[off + j] : ~0ull;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t j = tid + q * kSelectThreads;
        if (j < nb2) {
            uint64_t v = ~0ull;
            if (j < band_samples) {
                const uint32_t id = static_cast<uint32_t>(held[q]);
                const uint32_t r = id / w;
                const uint32_t key = sector_key(r, id - r * w, ax.s_sight, ax.c_sight);
                v = (static_cast<uint64_t>(key) << 32) | id;
            }
            cand[j] = v;
        }
    }
    __syncthreads();
    bitonic_sort(cand, nb2);

    const uint32_t middle = band_samples / 2u;
    const uint32_t end_of_band = band_size * 2u;
    const uint32_t n_steps = band_size - 1u;
    int32_t *deltas = deltas_base + static_cast<size_t>(sector) * stride;
    float *dists = dists_base + static_cast<size_t>(sector) * stride;
    const float pov_distance =
        static_cast<float>(rotated_distance(pov_id, w, ax.s_sight, ax.c_sight));
    for (uint32_t i = tid; i < n_steps; i += kSelectThreads) {
        int32_t delta = 0;
        float distance = 0.0f;
        const uint32_t band_id = middle + i;
        if (band_id < end_of_band) {
            const uint32_t cur = static_cast<uint32_t>(cand[band_id]);
            const uint32_t next = static_cast<uint32_t>(cand[band_id + 1]);
            delta = static_cast<int32_t>(cur) - static_cast<int32_t>(next);
            distance = fabsf(pov_distance -
                             static_cast<float>(rotated_distance(next, w, ax.s_sight, ax.c_sight)));
        }
        deltas[i] = delta;
        dists[i] = distance;
    }
    if (tid == 0) status[sector] = 0u;
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
        hipLaunchKernelGGL(sort_rowscan_kernel, dim3(256), dim3(kScanThreads), 0, c->stream,
                           s.hist, tiles, s.hist + 256u * tiles);
        hipLaunchKernelGGL(sort_scatter_kernel, dim3(tiles), dim3(kSortThreads), 0,
                           c->stream, ks, is, kd, id, static_cast<uint32_t>(n), shift,
                           s.hist, tiles, s.hist + 256u * tiles);
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
                       c->tab, first_slot, c->geo.band_deltas_size,
                       static_cast<uint32_t>(dem_copy_offset(c->geo.dem_size)),
                       pair_parity(c->geo.dem_width, c->geo.max_los_as_points));
    return hipGetLastError();
}

uint32_t max_band_samples() { return kMaxBandSamples; }

// Band data of sectors [first, first+count) by selection (no full sorts).
// `status[s]` is set non-zero for a sector whose window plus ties did not fit
// the LDS candidate buffer; the caller computes those with the sorts.
hipError_t launch_band_select(tvs_ctx *c, uint32_t first, uint32_t count,
                              const void *d_axes, uint32_t *d_status) {
    if (count == 0) return hipSuccess;
    const tvs_geometry &g = c->geo;
    hipLaunchKernelGGL(band_select_kernel, dim3(count), dim3(kSelectThreads), 0, c->stream,
                       static_cast<const SectorAxes *>(d_axes), first, g.dem_width,
                       g.band_size, g.band_samples, c->tab.delta, c->tab.dist,
                       c->tab.stride, d_status);
    return hipGetLastError();
}

uint32_t band_select_capacity() { return kSelectCapacity; }

}  // namespace tvs
