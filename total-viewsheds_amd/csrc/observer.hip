// observer.hip -- one observer's bands, one wavefront per band
// (kernel.rs:71-256 for the two kernel ids of one observer in each sector).
//
// The lane-per-band walks are the fast way to walk every band of a sector;
// for the bands of ONE observer they would walk all 2N bands to keep two.
// Here each band is one wave: the 64 lanes take 64 consecutive steps of the
// band at a time.  Per chunk of 64 steps:
//
//  * every lane computes its step's angle exactly as the reference does
//    (IEEE divide, kernel.rs:175), from the band's cumulative offsets;
//  * the running maximum before each step is an exclusive prefix maximum
//    across the lanes (maxNum, so a NaN angle is never the maximum) with
//    the previous chunks' maximum carried in.  `max = v ? angle : max` with
//    v = angle > max keeps, step after step, the maxNum of -2000 and every
//    angle so far (equal values differ at most in the sign of zero, which no
//    comparison sees), so step k is visible exactly when its angle exceeds
//    that prefix maximum;
//  * a ring event is written wherever the visibility changes (kernel.rs:
//    194-229; opening and closing are exclusive); its slot is the number of
//    events so far plus the events of the lower lanes (ballot + mbcnt);
//  * the band surface is the reference's sequential f32 sum of d x 0.0174533
//    over the visible steps in step order (kernel.rs:216): a wave-uniform
//    loop over the ballot of visible lanes.
//
// After the last step a visible band gets its final close (kernel.rs:237-
// 240), and the count slot is 1 + events (kernel.rs:243).
#include <algorithm>
#include <string>
#include <vector>

#include "engine.hpp"

namespace {

constexpr uint32_t kWave = 64;

// v of the lane d below (a lane's own v below lane d).  All 64 lanes must
// be active when it runs.
__device__ __forceinline__ float shfl_up_f(float v, int d) {
    const int lane = static_cast<int>(__lane_id());
    const int src = lane >= d ? lane - d : lane;
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

// Block b walks band (sector slot first + b / 2, forward if b is even).
// counts, surfaces: one entry per band; events: `cap` slots per band.
__global__ __launch_bounds__(kWave) void observer_walk_kernel(
    const float *__restrict__ elev, const int32_t *__restrict__ off,
    const float *__restrict__ dist, const float *__restrict__ curv,
    const float *__restrict__ wt, uint32_t stride, uint32_t first, uint32_t n_steps,
    uint32_t pov, float obs, uint32_t *__restrict__ counts, float *__restrict__ surfaces,
    uint32_t *__restrict__ events, uint32_t cap) {
    const uint32_t band = blockIdx.x;
    const uint32_t slot = first + (band >> 1);
    const bool backward = (band & 1u) != 0;
    const uint32_t lane = threadIdx.x;
    const size_t tb = static_cast<size_t>(slot) * stride;
    const float h0 = elev[pov] + obs;                                    // kernel.rs:134
    uint32_t *band_events = events + static_cast<size_t>(band) * cap;
    float carry = tvs::kMaxAngle;       // the maximum before the chunk (kernel.rs:102)
    bool prev_visible = true;           // kernel.rs:146
    float surface = 0.0f;
    uint32_t n_events = 0;
    uint32_t last_id = pov;
    bool last_visible = true;
    for (uint32_t k0 = 0; k0 < n_steps; k0 += kWave) {
        const uint32_t k = k0 + lane;
        const bool valid = k < n_steps;
        float angle = -INFINITY;        // spare lanes: never visible, never the maximum
        uint32_t id = pov;
        if (valid) {
            const int32_t o = off[tb + k];
            id = backward ? pov - static_cast<uint32_t>(o) : pov + static_cast<uint32_t>(o);
            const float d = dist[tb + k];
            angle = (elev[id] - h0) / d - curv[tb + k];                 // kernel.rs:175
        }
        // Inclusive prefix maximum across the lanes, then exclusive.
        float incl = angle;
#pragma unroll
        for (int s = 1; s < static_cast<int>(kWave); s <<= 1) {
            const float other = shfl_up_f(incl, s);
            if (static_cast<int>(lane) >= s) incl = fmaxf(incl, other);
        }
        // The shuffle runs on every lane: a lane reading a disabled lane's
        // register through ds_bpermute gets 0, not its value.
        const float up = shfl_up_f(incl, 1);
        const float before = lane == 0 ? -INFINITY : up;
        const float max_before = fmaxf(carry, before);
        const bool visible = valid && angle > max_before;               // kernel.rs:190
        const uint64_t vis_mask = __ballot(visible);
        // Visibility of the step before each lane's step.
        const bool prev = lane == 0 ? prev_visible : ((vis_mask >> (lane - 1)) & 1ull) != 0;
        const bool event = valid && visible != prev;
        const uint64_t ev_mask = __ballot(event);
        if (event) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                static_cast<uint32_t>(ev_mask >> 32),
                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ev_mask), 0u));
            band_events[n_events + below] = id;
        }
        n_events += static_cast<uint32_t>(__popcll(ev_mask));
        // Surface: the visible steps' weights added in step order.
        for (uint64_t m = vis_mask; m; m &= m - 1) {
            const uint32_t j = static_cast<uint32_t>(__builtin_ctzll(m));
            surface = surface + wt[tb + k0 + j];                        // kernel.rs:216
        }
        const uint32_t last = (n_steps - k0 < kWave ? n_steps - k0 : kWave) - 1u;
        carry = fmaxf(carry, __shfl(incl, static_cast<int>(last)));
        prev_visible = ((vis_mask >> last) & 1ull) != 0;
        last_visible = prev_visible;
        last_id = static_cast<uint32_t>(__shfl(static_cast<int>(id), static_cast<int>(last)));
    }
    if (last_visible) {                                                 // kernel.rs:237-240
        if (lane == 0) band_events[n_events] = last_id;
        ++n_events;
    }
    if (lane == 0) {
        counts[band] = 1u + n_events;                                   // kernel.rs:243
        surfaces[band] = surface;
    }
}

tvs_status ofail(tvs_ctx *c, tvs_status st, const std::string &msg) {
    c->err = msg;
    return st;
}

#define TVS_OHIP(c, expr, what)                                                    \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return ofail((c), _e == hipErrorOutOfMemory ? TVS_ERR_OUT_OF_MEMORY    \
                                                        : TVS_ERR_HIP,             \
                         std::string(what) + ": " + hipGetErrorString(_e));         \
    } while (0)

}  // namespace

extern "C" {

tvs_status tvs_observer_walk(tvs_ctx *c, uint32_t tvs_id, uint32_t first, uint32_t count,
                             uint32_t *counts, float *band_surfaces, uint32_t *events,
                             size_t events_capacity, uint64_t *n_events) {
    tvs::TraceRange trace_range("tvs_observer_walk");
    if (!c) return TVS_ERR_INVALID_ARGUMENT;
    if (first >= tvs::kSectorSteps || count > tvs::kSectorSteps - first)
        return ofail(c, TVS_ERR_INVALID_ARGUMENT, "sector range outside 0..180");
    const tvs_geometry &g = c->geo;
    if (tvs_id >= g.computable_points_count)
        return ofail(c, TVS_ERR_INVALID_ARGUMENT, "observer index beyond the computable points");
    if (!c->has_f32) return ofail(c, TVS_ERR_STATE, "no elevations uploaded");
    for (uint32_t s = first; s < first + count; ++s)
        if (!c->band_ready[s])
            return ofail(c, TVS_ERR_STATE, "sector " + std::to_string(s) + " has no band data");
    if (!counts) return ofail(c, TVS_ERR_INVALID_ARGUMENT, "null counts");
    if (count == 0) {
        if (n_events) *n_events = 0;
        return TVS_OK;
    }
    const uint32_t T = g.tvs_width, L = g.max_los_as_points, w = g.dem_width;
    const uint32_t pov = (tvs_id / T + L) * w + (tvs_id % T + L);      // kernel.rs:125-131
    const uint32_t bands = 2u * count;
    const uint32_t cap = g.band_deltas_size + 1u;   // at most one event per step, plus the close
    const size_t need = static_cast<size_t>(bands) * (cap + 2u);
    if (c->obs_cap < need) {
        if (c->d_obs) (void)hipFree(c->d_obs);
        c->d_obs = nullptr;
        c->obs_cap = 0;
        TVS_OHIP(c, hipMalloc(reinterpret_cast<void **>(&c->d_obs), need * sizeof(uint32_t)),
                 "observer walk buffers");
        c->obs_cap = need;
    }
    uint32_t *d_counts = c->d_obs;
    float *d_surf = reinterpret_cast<float *>(c->d_obs + bands);
    uint32_t *d_events = c->d_obs + 2u * bands;
    hipLaunchKernelGGL(observer_walk_kernel, dim3(bands), dim3(kWave), 0, c->stream,
                       c->elev_f32, c->tab.off, c->tab.dist, c->tab.curv, c->tab.wt,
                       c->tab.stride, first, g.band_deltas_size, pov,
                       c->params.observer_height, d_counts, d_surf, d_events, cap);
    TVS_OHIP(c, hipGetLastError(), "observer walk");
    std::vector<uint32_t> h_counts(bands);
    std::vector<float> h_surf(bands);
    TVS_OHIP(c, hipMemcpyAsync(h_counts.data(), d_counts, bands * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, c->stream), "observer counts");
    TVS_OHIP(c, hipMemcpyAsync(h_surf.data(), d_surf, bands * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream), "observer surfaces");
    TVS_OHIP(c, hipStreamSynchronize(c->stream), "observer walk");
    uint64_t total = 0;
    for (uint32_t b = 0; b < bands; ++b) total += h_counts[b] - 1u;
    std::copy(h_counts.begin(), h_counts.end(), counts);
    if (band_surfaces) std::copy(h_surf.begin(), h_surf.end(), band_surfaces);
    if (n_events) *n_events = total;
    if (!events) return TVS_OK;
    if (events_capacity < total)
        return ofail(c, TVS_ERR_SIZE, "events buffer holds " + std::to_string(events_capacity) +
                                          ", the observer has " + std::to_string(total));
    std::vector<uint32_t> slots(static_cast<size_t>(bands) * cap);
    TVS_OHIP(c, hipMemcpy(slots.data(), d_events, slots.size() * sizeof(uint32_t),
                          hipMemcpyDeviceToHost), "observer events");
    uint64_t at = 0;
    for (uint32_t b = 0; b < bands; ++b) {
        const uint32_t n = h_counts[b] - 1u;
        std::copy(slots.begin() + static_cast<size_t>(b) * cap,
                  slots.begin() + static_cast<size_t>(b) * cap + n, events + at);
        at += n;
    }
    return TVS_OK;
}

}  // extern "C"
