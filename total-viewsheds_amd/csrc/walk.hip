// walk.hip -- the band-of-sight walk (kernel.rs:71-256) on CDNA4.
//
// Two implementations of the same per-band loop:
//
//  * walk_plain_kernel: one sector per launch, one thread per band, the
//    reference loop as written (IEEE f32 divide per step, cumulative DEM id,
//    per-step band data read from memory).  Threads [0,128) of a block walk
//    the forward bands of 128 consecutive points and threads [128,256) their
//    backward bands; the block then adds forward-then-backward into the
//    cumulative surface, the order compute.rs:229 + kernel.rs:254 produce.
//
//  * walk_batched_kernel: every requested sector in one launch.  A work item
//    is (sector, direction, 256 consecutive points); lanes are consecutive
//    points of a row, so each step's gather is one contiguous row segment at
//    a wave-uniform offset (scalar base + per-lane constant byte offset: no
//    per-step address arithmetic in the vector unit).  The per-step constants
//    (cumulative offset, distance, 1/distance, curvature term, surface
//    weight) come from the sector's step table through the scalar cache.
//    Each band's surface goes to a per-(sector, direction, point) slot; a
//    reduction then adds them per point in sector order, forward before
//    backward, giving the reference's f32 cumulative sum bit for bit.
#include <algorithm>

#include "engine.hpp"
#include "fastdiv.hpp"

namespace tvs {
namespace {

constexpr uint32_t kBlock = 256;
constexpr int kUnroll = 8;

__global__ __launch_bounds__(kBlock) void walk_plain_kernel(
    const float *__restrict__ elev, const int32_t *__restrict__ deltas,
    const float *__restrict__ dists, uint32_t n_steps, uint32_t N, uint32_t T,
    uint32_t L, uint32_t w, float obs, float *__restrict__ cum,
    uint32_t *__restrict__ ring, uint32_t R,
    unsigned long long *__restrict__ overflow) {
    __shared__ float forward_surface[kBlock / 2];
    const uint32_t lt = threadIdx.x & (kBlock / 2 - 1);
    const bool backward = threadIdx.x >= kBlock / 2;
    const uint32_t tvs_id = blockIdx.x * (kBlock / 2) + lt;
    float band_surface = 0.0f;
    if (tvs_id < N) {
        const uint32_t kernel_id = backward ? tvs_id + N : tvs_id;   // 117-123
        const uint32_t pov_x = tvs_id % T + L;                        // 125
        const uint32_t pov_y = tvs_id / T + L;                        // 126
        const uint32_t pov_id = pov_y * w + pov_x;                    // 131
        const float pov_elevation = elev[pov_id] + obs;               // 134
        float max_angle = kMaxAngle;
        bool visible = true, previously_visible = true, closing = false;
        uint32_t ring_id = 1;
        uint32_t *band_ring = ring ? ring + static_cast<size_t>(kernel_id) * R : nullptr;
        unsigned dropped = 0;
        auto save = [&](uint32_t slot, uint32_t value) {             // 302-309
            if (!band_ring) return;
            if (slot < R) band_ring[slot] = value; else ++dropped;
        };
        save(ring_id, pov_id);                                       // 145
        uint32_t dem_id = pov_id;
        for (uint32_t i = 0; i < n_steps; ++i) {                     // 153
            const int32_t delta = deltas[i];
            dem_id = backward ? dem_id - static_cast<uint32_t>(delta)
                              : dem_id + static_cast<uint32_t>(delta);
            const float elevation_delta = elev[dem_id] - pov_elevation;
            const float distance = dists[i];
            const float angle = (elevation_delta / distance) -
                                (distance / kEarthRadiusSquared);    // 175
            visible = angle > max_angle;                              // 190
            const bool opening = visible && !previously_visible;
            closing = previously_visible && !visible;
            if (visible) band_surface += distance * kTanOneRad;       // 216
            if (opening) { ++ring_id; save(ring_id, dem_id); }
            if (closing) { ++ring_id; save(ring_id, dem_id); }
            previously_visible = visible;
            max_angle = fmaxf(angle, max_angle);                      // 233
        }
        if (visible && !closing) { ++ring_id; save(ring_id, dem_id); }  // 237-240
        if (band_ring && R > 0) band_ring[0] = ring_id;              // 243
        if (dropped) atomicAdd(overflow, static_cast<unsigned long long>(dropped));
    }
    if (!backward) forward_surface[lt] = band_surface;
    __syncthreads();
    if (backward && tvs_id < N) {                                    // 254
        float c = cum[tvs_id];
        c = c + forward_surface[lt];
        c = c + band_surface;
        cum[tvs_id] = c;
    }
}

// Heights are read through a buffer resource whose base sits `bias`
// elements below the DEM: the wave-uniform step offset (made non-negative
// by the bias) is the scalar offset of the load and the lane's point is its
// vector offset, so a step costs no vector address arithmetic.  The
// resource's range covers exactly [elev - bias, elev + w*w).
template <typename H>
__device__ __forceinline__ float load_height(__amdgpu_buffer_rsrc_t rsrc,
                                             uint32_t pov_bytes,
                                             uint32_t soff_bytes);

template <>
__device__ __forceinline__ float load_height<int16_t>(__amdgpu_buffer_rsrc_t rsrc,
                                                      uint32_t pov_bytes,
                                                      uint32_t soff_bytes) {
    const int16_t h = static_cast<int16_t>(
        __builtin_amdgcn_raw_buffer_load_b16(rsrc, pov_bytes, soff_bytes, 0));
    return static_cast<float>(h);
}

template <>
__device__ __forceinline__ float load_height<float>(__amdgpu_buffer_rsrc_t rsrc,
                                                    uint32_t pov_bytes,
                                                    uint32_t soff_bytes) {
    return __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(rsrc, pov_bytes, soff_bytes, 0));
}

// Buffer resource word 3 for raw (untyped) loads on gfx9-family parts.
constexpr int kBufferFlags = 0x00020000;

struct BatchedArgs {
    const int32_t *off;
    const float *dist;
    const float *rcp;
    const float *rcp_lo;
    const float *curv;
    const float *wt;
    const SectorMeta *meta;
    uint32_t stride;
    uint32_t n_steps;
    uint32_t first_slot;
    uint32_t N, T, L, w;
    float obs;
    int32_t bias;               // elements; >= every |offset| of the run
    uint32_t blocks_per_pass;
    uint32_t n_items;
    uint32_t items_per_xcd;
    float *partial;
    uint32_t *ring;
    uint32_t R;
    uint32_t capture_slots;
    unsigned long long *overflow;
};

template <typename H, bool FAST, bool CAPTURE>
__global__ __launch_bounds__(kBlock) void walk_batched_kernel(
    const H *__restrict__ elev, const BatchedArgs a) {
    // Consecutive block ids land on the 8 XCDs round-robin; give each XCD a
    // contiguous run of work items so its L2 holds one region of the DEM.
    const uint32_t b = blockIdx.x;
    const uint32_t item = (b & 7u) * a.items_per_xcd + (b >> 3);
    if (item >= a.n_items) return;
    const uint32_t pass = item / a.blocks_per_pass;      // 2*sector + dir
    const uint32_t ob = item - pass * a.blocks_per_pass;
    const uint32_t rel = pass >> 1;
    const bool backward = (pass & 1u) != 0;
    const uint32_t slot = a.first_slot + rel;

    const uint32_t t_raw = ob * kBlock + threadIdx.x;
    const bool active = t_raw < a.N;
    const uint32_t t = active ? t_raw : a.N - 1u;       // spare lanes shadow
    const uint32_t pov = (t / a.T + a.L) * a.w + (t % a.T + a.L);
    const uint32_t pov_bytes = pov * static_cast<uint32_t>(sizeof(H));
    const float h0 = static_cast<float>(elev[pov]) + a.obs;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<H *>(elev - a.bias), 0,
        static_cast<int>((static_cast<uint32_t>(a.bias) + a.w * a.w) * sizeof(H)),
        kBufferFlags);
    const int32_t sgn = backward ? -1 : 1;
    // Scalar byte offset of a step: (sgn * off + bias) * sizeof(H) >= 0.
    auto soff = [&](int32_t o) -> uint32_t {
        return static_cast<uint32_t>(sgn * o + a.bias) * static_cast<uint32_t>(sizeof(H));
    };

    const size_t tb = static_cast<size_t>(slot) * a.stride;
    const int32_t *__restrict__ off =
        static_cast<const int32_t *>(__builtin_assume_aligned(a.off + tb, 256));
    const float *__restrict__ dist =
        static_cast<const float *>(__builtin_assume_aligned(a.dist + tb, 256));
    const float *__restrict__ rcp =
        static_cast<const float *>(__builtin_assume_aligned(a.rcp + tb, 256));
    const float *__restrict__ rcp_lo =
        static_cast<const float *>(__builtin_assume_aligned(a.rcp_lo + tb, 256));
    const float *__restrict__ curv =
        static_cast<const float *>(__builtin_assume_aligned(a.curv + tb, 256));
    const float *__restrict__ wt =
        static_cast<const float *>(__builtin_assume_aligned(a.wt + tb, 256));
    const uint32_t n_fast = a.meta[slot].n_fast;
    const uint32_t tail_start = a.meta[slot].tail_start;

    // `max_angle = v ? angle : max_angle` is f32::max(angle, max_angle)
    // (kernel.rs:233) for every input the walk meets: v is `angle >
    // max_angle`, so a NaN angle (false) keeps the max as maxNum does.
    float max_angle = kMaxAngle;
    float surface = 0.0f;
    bool visible = true;

    // Ring records (kernel.rs:141-243): an event is written exactly when the
    // visibility changes (opening and closing are exclusive), plus a final
    // close if the last step is visible.
    const bool capture = CAPTURE && active && rel < a.capture_slots;
    uint32_t *band_ring = nullptr;
    uint32_t ring_id = 1;
    unsigned dropped = 0;
    if (CAPTURE && capture) {
        const uint32_t kernel_id = backward ? t + a.N : t;
        band_ring = a.ring + static_cast<size_t>(rel) * 2u * a.N * a.R +
                    static_cast<size_t>(kernel_id) * a.R;
        if (1u < a.R) band_ring[1] = pov; else ++dropped;
    }
    auto event = [&](int32_t o) {
        ++ring_id;
        if (ring_id < a.R) band_ring[ring_id] = pov + static_cast<uint32_t>(sgn * o);
        else ++dropped;
    };

    uint32_t k = 0;
    const uint32_t fast_end = n_fast - n_fast % kUnroll;
    for (; k < fast_end; k += kUnroll) {
        float h[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) h[u] = load_height<H>(rsrc, pov_bytes, soff(off[k + u]));
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const float dh = h[u] - h0;
            const float q = FAST ? div_by_reciprocal(dh, dist[k + u], rcp[k + u], rcp_lo[k + u])
                                 : dh / dist[k + u];
            const float angle = q - curv[k + u];
            const bool v = angle > max_angle;
            if (CAPTURE && capture && v != visible) event(off[k + u]);
            const float grown = surface + wt[k + u];
            surface = v ? grown : surface;
            max_angle = v ? angle : max_angle;
            visible = v;
        }
    }
    for (; k < tail_start; ++k) {
        const int32_t o = off[k];
        const float d = dist[k];
        const float dh = load_height<H>(rsrc, pov_bytes, soff(o)) - h0;
        const float q = (FAST && k < n_fast) ? div_by_reciprocal(dh, d, rcp[k], rcp_lo[k]) : dh / d;
        const float angle = q - curv[k];
        const bool v = angle > max_angle;
        if (CAPTURE && capture && v != visible) event(o);
        const float grown = surface + wt[k];
        surface = v ? grown : surface;
        max_angle = v ? angle : max_angle;
        visible = v;
    }
    if (k < a.n_steps) {
        // Zero tail: delta 0 keeps the point, distance +0 gives the same
        // angle (dh/0 - 0: +inf, -inf or NaN) at every step, and the weight
        // 0 * 0.0174533 leaves the surface unchanged.
        const int32_t o = off[k];
        const float dh = load_height<H>(rsrc, pov_bytes, soff(o)) - h0;
        const float angle = dh / dist[k] - curv[k];
        for (; k < a.n_steps; ++k) {
            const bool v = angle > max_angle;
            if (CAPTURE && capture && v != visible) event(o);
            max_angle = v ? angle : max_angle;
            visible = v;
        }
    }
    if (CAPTURE && capture) {
        if (visible) event(a.n_steps ? off[a.n_steps - 1] : 0);
        if (a.R > 0) band_ring[0] = ring_id;
        if (dropped) atomicAdd(a.overflow, static_cast<unsigned long long>(dropped));
    }
    if (active) a.partial[static_cast<size_t>(pass) * a.N + t] = surface;
}

// Per point, in sector order: cum += forward; cum += backward (the f32
// order of the reference's serial loop), and the same sums in f64.
__global__ __launch_bounds__(kBlock) void reduce_kernel(
    const float *__restrict__ partial, uint32_t count, uint32_t N,
    float *__restrict__ cum, double *__restrict__ cum64) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= N) return;
    float c = cum[t];
    double s = cum64[t];
    for (uint32_t r = 0; r < count; ++r) {
        const float f = partial[static_cast<size_t>(2u * r) * N + t];
        const float b = partial[static_cast<size_t>(2u * r + 1u) * N + t];
        c = c + f;
        c = c + b;
        s = s + static_cast<double>(f);
        s = s + static_cast<double>(b);
    }
    cum[t] = c;
    cum64[t] = s;
}

__global__ void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void widen_i16_kernel(const int16_t *in, float *out, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = static_cast<float>(in[i]);
}

template <typename H, bool FAST, bool CAPTURE>
hipError_t launch_batched_t(tvs_ctx *c, const H *elev, const BatchedArgs &a,
                            uint32_t grid) {
    hipLaunchKernelGGL((walk_batched_kernel<H, FAST, CAPTURE>), dim3(grid),
                       dim3(kBlock), 0, c->stream, elev, a);
    return hipGetLastError();
}

template <typename H>
hipError_t launch_batched_h(tvs_ctx *c, const H *elev, const BatchedArgs &a,
                            uint32_t grid, bool fast, bool capture) {
    if (fast) {
        return capture ? launch_batched_t<H, true, true>(c, elev, a, grid)
                       : launch_batched_t<H, true, false>(c, elev, a, grid);
    }
    return capture ? launch_batched_t<H, false, true>(c, elev, a, grid)
                   : launch_batched_t<H, false, false>(c, elev, a, grid);
}

}  // namespace

hipError_t launch_walk_plain(tvs_ctx *c, uint32_t slot, uint32_t *ring_out) {
    const tvs_geometry &g = c->geo;
    const uint32_t N = g.computable_points_count;
    const uint32_t grid = (N + kBlock / 2 - 1) / (kBlock / 2);
    const size_t tb = static_cast<size_t>(slot) * c->tab.stride;
    hipLaunchKernelGGL(walk_plain_kernel, dim3(grid), dim3(kBlock), 0, c->stream,
                       c->elev_f32, c->tab.delta + tb, c->tab.dist + tb,
                       g.band_deltas_size, N, g.tvs_width, g.max_los_as_points,
                       g.dem_width, c->params.observer_height, c->cum, ring_out,
                       c->rings_per_band, c->overflow);
    return hipGetLastError();
}

hipError_t launch_walk_batched(tvs_ctx *c, uint32_t first_slot, uint32_t count,
                               uint32_t capture_slots) {
    const tvs_geometry &g = c->geo;
    BatchedArgs a;
    a.off = c->tab.off;
    a.dist = c->tab.dist;
    a.rcp = c->tab.rcp;
    a.rcp_lo = c->tab.rcp_lo;
    a.curv = c->tab.curv;
    a.wt = c->tab.wt;
    a.meta = c->tab.meta;
    a.stride = c->tab.stride;
    a.n_steps = g.band_deltas_size;
    a.first_slot = first_slot;
    a.N = g.computable_points_count;
    a.T = g.tvs_width;
    a.L = g.max_los_as_points;
    a.w = g.dem_width;
    a.obs = c->params.observer_height;
    int32_t bias = 0;
    for (uint32_t s = first_slot; s < first_slot + count; ++s)
        bias = std::max(bias, std::max(c->meta_host[s].max_off, -c->meta_host[s].min_off));
    a.bias = bias;
    a.blocks_per_pass = (a.N + kBlock - 1) / kBlock;
    a.n_items = a.blocks_per_pass * 2u * count;
    a.items_per_xcd = (a.n_items + 7u) / 8u;
    a.partial = c->partial;
    a.ring = c->ring;
    a.R = c->rings_per_band;
    a.capture_slots = capture_slots;
    a.overflow = c->overflow;
    const uint32_t grid = a.items_per_xcd * 8u;
    const bool capture = capture_slots > 0 && c->ring != nullptr;
    if (c->has_i16)
        return launch_batched_h<int16_t>(c, c->elev_i16, a, grid, c->fast_div_ok, capture);
    return launch_batched_h<float>(c, c->elev_f32, a, grid, c->fast_div_ok, capture);
}

hipError_t launch_reduce(tvs_ctx *c, uint32_t count) {
    const uint32_t N = c->geo.computable_points_count;
    hipLaunchKernelGGL(reduce_kernel, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       c->stream, c->partial, count, N, c->cum, c->cum64);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                       0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_widen_i16(hipStream_t s, const int16_t *in, float *out, size_t n) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(widen_i16_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                       0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace tvs
