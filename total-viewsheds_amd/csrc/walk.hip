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
//  * walk_tiles_kernel: every requested sector of one or many DEM tiles in
//    one launch.  A work item is (tile, sector, direction, BLOCK consecutive
//    points); lanes are consecutive points of a row, so each step's gather is
//    one contiguous row segment at a wave-uniform offset (scalar base +
//    per-lane constant byte offset: no per-step address arithmetic in the
//    vector unit).  The per-step constants (cumulative offset, distance,
//    two-term reciprocal, curvature term, surface weight) come from the
//    sector's step table through the scalar cache.  Each band's surface goes
//    to a per-(sector, direction, point) slot; reduce_tiles_kernel then adds
//    them per point in sector order, forward before backward, giving the
//    reference's f32 cumulative sum bit for bit.
#include <algorithm>

#include "engine.hpp"
#include "fastdiv.hpp"

// Typed buffer load (MUBUF buffer_load_format_x), declared by its LLVM
// intrinsic name: the texture unit converts the element with the
// resource's data/number format.
__device__ float tvs_raw_buffer_load_format_f32(__amdgpu_buffer_rsrc_t rsrc, int voffset,
                                                int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.format.f32");

namespace tvs {
namespace {

constexpr uint32_t kBlock = 256;
// Waves per SIMD the int16 eight-band walk (two pairs per lane, both
// directions) is compiled for: left alone it takes ~98 VGPRs (four waves);
// held to five waves it fits in 81 with no spill and runs ~2-3 % faster
// than the four-band walk on 1201 x 1201 (same box), where with six it only
// ties and with seven it spills.  The f32 form (two dword loads per pair)
// spills at five and is left alone.
constexpr int kQuadWaves = 5;

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
// elements below the tile: the wave-uniform step offset (made non-negative
// by the bias, and moved into the DEM's second copy where a two-point load
// would be misaligned in the first) is the scalar offset of the load and
// the lane's point is its vector offset.  The resource's range covers
// exactly [elev - bias, elev + span): both copies.
template <typename H>
__device__ __forceinline__ float load_height(__amdgpu_buffer_rsrc_t rsrc,
                                             uint32_t pov_bytes,
                                             uint32_t soff_bytes);

// int16 heights: the resource's format (16-bit, SSCALED) makes the load
// return float(h) exactly, so no conversion instruction is spent per step.
template <>
__device__ __forceinline__ float load_height<int16_t>(__amdgpu_buffer_rsrc_t rsrc,
                                                      uint32_t pov_bytes,
                                                      uint32_t soff_bytes) {
    return tvs_raw_buffer_load_format_f32(rsrc, static_cast<int>(pov_bytes),
                                          static_cast<int>(soff_bytes), 0);
}

template <>
__device__ __forceinline__ float load_height<float>(__amdgpu_buffer_rsrc_t rsrc,
                                                    uint32_t pov_bytes,
                                                    uint32_t soff_bytes) {
    return __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(rsrc, pov_bytes, soff_bytes, 0));
}

// surface + w when v, else surface: fma(1, w, s) = RN(s + w) and
// fma(0, w, s) = s for the finite w >= +0 and s != -0 of the walk
// (kernel.rs:216).  Written as one unconditional FMA so the compiler keeps
// the step branch-free instead of sinking the weight's load into a branch.
__device__ __forceinline__ float add_if(bool v, float surface, float w) {
    return __builtin_fmaf(v ? 1.0f : 0.0f, w, surface);
}

// Scalar byte offset of a step for heights of type H, from its int16 byte
// position in the step table.
template <typename H>
__device__ __forceinline__ uint32_t soff(uint32_t pos16) {
    return sizeof(H) == 2 ? pos16 : pos16 << 1;
}

// Buffer resource word 3 (gfx9 family): DST_SEL_X [2:0], NUM_FORMAT
// [14:12], DATA_FORMAT [18:15].  Untyped loads need a non-zero data format
// (32); the int16 heights use DST_SEL_X = X (4), NUM_FORMAT = SSCALED (3),
// DATA_FORMAT = 16 (2).
template <typename H> constexpr int kHeightRsrcFlags = 0x00020000;
template <> constexpr int kHeightRsrcFlags<int16_t> = 4 | (3 << 12) | (2 << 15);

// One step's update where the walk records no rings (kernel.rs:216, 233):
// v_cmp finds the lanes with angle > max_angle; only if there is one does
// exec narrow to them, which add w to the surface and take the angle as
// their new maximum, and exec is then set back to `lanes`, the mask read at
// the start of the walk (the walk has no other divergent control flow, so
// exec equals it at every step).  A lane sees ~2 % of its steps and a wave
// of 64 lanes sees a band's point at ~35 % of the steps (1201 x 1201, 90 m),
// so the branch skips the masked add and move at two steps of three: 4-5 %
// faster than narrowing exec with v_cmpx at every step (same box, bit-
// identical).  The masked move is `max_angle = angle > max_angle ? angle :
// max_angle`, which is f32::max(angle, max_angle) for the comparisons that
// follow: a NaN angle compares false and leaves the maximum (as maxNum
// does), and values equal under > differ at most in the sign of zero, which
// no comparison sees.  A move issues alongside more kinds of vector
// instruction than v_max does (the pairing counters), which made the walk
// 3-4 % faster.
__device__ __forceinline__ void step_update(float angle, float w, uint64_t lanes,
                                            float &max_angle, float &surface) {
    asm("v_cmp_gt_f32 vcc, %[a], %[m]\n\t"
        "s_cbranch_vccz 1f\n\t"
        "s_mov_b64 exec, vcc\n\t"
        "v_add_f32 %[s], %[w], %[s]\n\t"
        "v_mov_b32 %[m], %[a]\n\t"
        "s_mov_b64 exec, %[e]\n"
        "1:"
        : [m] "+v"(max_angle), [s] "+v"(surface)
        : [a] "v"(angle), [w] "s"(w), [e] "s"(lanes)
        : "vcc");
}

// step_update for a lane's forward (f) and backward (b) band at the same
// step.
__device__ __forceinline__ void pair_update(float angle_f, float angle_b, float w,
                                            uint64_t lanes, float &max_f, float &max_b,
                                            float &surface_f, float &surface_b) {
    asm("v_cmp_gt_f32 vcc, %[af], %[mf]\n\t"
        "s_cbranch_vccz 1f\n\t"
        "s_mov_b64 exec, vcc\n\t"
        "v_add_f32 %[sf], %[w], %[sf]\n\t"
        "v_mov_b32 %[mf], %[af]\n\t"
        "s_mov_b64 exec, %[e]\n"
        "1:\n\t"
        "v_cmp_gt_f32 vcc, %[ab], %[mb]\n\t"
        "s_cbranch_vccz 2f\n\t"
        "s_mov_b64 exec, vcc\n\t"
        "v_add_f32 %[sb], %[w], %[sb]\n\t"
        "v_mov_b32 %[mb], %[ab]\n\t"
        "s_mov_b64 exec, %[e]\n"
        "2:"
        : [mf] "+v"(max_f), [mb] "+v"(max_b), [sf] "+v"(surface_f), [sb] "+v"(surface_b)
        : [af] "v"(angle_f), [ab] "v"(angle_b), [w] "s"(w), [e] "s"(lanes)
        : "vcc");
}

// step_update with the surface weight in a vector register.
__device__ __forceinline__ void step_update_v(float angle, float w, uint64_t lanes,
                                              float &max_angle, float &surface) {
    asm("v_cmp_gt_f32 vcc, %[a], %[m]\n\t"
        "s_cbranch_vccz 1f\n\t"
        "s_mov_b64 exec, vcc\n\t"
        "v_add_f32 %[s], %[w], %[s]\n\t"
        "v_mov_b32 %[m], %[a]\n\t"
        "s_mov_b64 exec, %[e]\n"
        "1:"
        : [m] "+v"(max_angle), [s] "+v"(surface)
        : [a] "v"(angle), [w] "v"(w), [e] "s"(lanes)
        : "vcc");
}

// A wave-uniform value copied into a vector register, so that the
// instructions reading it take no scalar operand (two instructions that
// each read a scalar register never issue together).
__device__ __forceinline__ float to_vgpr(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

// h - h0 from a dword of two f16 heights h16 = h - C (low or high half) and
// a lane's c = C - h0, exact in f32: one mixed-precision FMA, h16 * 1 + c,
// converts the half exactly and rounds h16 + c = h - h0 once -- the f32
// subtraction of the f32 height the walk would otherwise convert and
// subtract (kernel.rs:160-163), in one instruction instead of two.
__device__ __forceinline__ float half_lo_diff(uint32_t w, float c) {
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(c));
    return r;
}
__device__ __forceinline__ float half_hi_diff(uint32_t w, float c) {
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=v"(r) : "v"(w), "v"(c));
    return r;
}

// c = C - h0 and whether it is exact (the rounding error of the f32
// subtraction, by Knuth's TwoSum, is zero).
__device__ __forceinline__ bool exact_offset(float C, float h0, float &c) {
    c = C - h0;
    const float bb = c - C;
    const float err = (C - (c - bb)) + (-h0 - bb);
    return err == 0.0f;
}

// Pointers read out of a TileDesc are generic; naming them global lets the
// wave-uniform table reads become scalar loads (s_load_dwordx8) instead of
// per-lane loads plus a readfirstlane loop for the scalar offset.
#define TVS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const TVS_GLOBAL T *global_ptr(const T *p) {
    return (const TVS_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ TVS_GLOBAL T *global_ptr(T *p) {
    return (TVS_GLOBAL T *)p;
}

// The step tables and sector metadata are not written while a walk runs;
// reading them through the constant address space lets every wave-uniform
// read be a scalar load whatever stores the kernel makes elsewhere (with
// global pointers the compiler must prove no store can alias the table,
// and in the larger kernels it stops proving it and falls back to per-lane
// vector loads, which compete with the height loads).
#define TVS_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const TVS_CONST T *const_ptr(const T *p) {
    return (const TVS_CONST T *)p;
}

// The divide of a regular step.  DIV 4: the split-reciprocal divide
// (proven correctly rounded in the fast-divide range); DIV 2: the short
// divide (used only for sectors verified against DIV 4); DIV 0: IEEE.
template <int DIV>
__device__ __forceinline__ float step_div(float a, float d, float y_hi, float y_lo) {
    if (DIV == 4) return div_by_reciprocal(a, d, y_hi, y_lo);
    if (DIV == 2) return div_short(a, y_hi, y_lo);
    return a / d;
}

// One band per lane: (tile d, sector slot d.first_slot + rel, direction)
// for point t.  DIV picks the divide of the regular steps (see step_div).
template <typename H, int DIV, bool CAPTURE, int UNROLL, bool PIPE>
__device__ __forceinline__ void walk_band(const TileDesc &d, uint32_t rel, bool backward,
                                          uint32_t t, bool active) {
    const H *elev = static_cast<const H *>(d.elev);
    const uint32_t slot = d.first_slot + rel;
    const uint32_t pov = (t / d.T + d.L) * d.w + (t % d.T + d.L);
    const uint32_t pov_bytes = (pov - d.pov_shift) * static_cast<uint32_t>(sizeof(H));
    const float h0 = static_cast<float>(global_ptr(elev)[pov]) + d.obs;
    const uint32_t head = const_ptr(d.meta)[slot].head;
    const uint32_t n_fast = const_ptr(d.meta)[slot].n_fast;
    const uint32_t tail_start = const_ptr(d.meta)[slot].tail_start;
    const uint32_t bias = const_ptr(d.meta)[slot].bias;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<H *>(elev - bias), 0,
        static_cast<int>((bias + d.span) * sizeof(H)), kHeightRsrcFlags<H>);

    const size_t tb = static_cast<size_t>(slot) * d.stride;
    // Position of each step relative to the resource base in int16 bytes:
    // 2 (bias + off) forward, 2 (bias - off) backward (from the step table).
    const TVS_CONST uint32_t *__restrict__ pos = const_ptr((backward ? d.pos_b : d.pos_f) + tb);
    const uint64_t lanes = __builtin_amdgcn_read_exec();
    const TVS_CONST float *__restrict__ dist = const_ptr(d.dist + tb);
    const TVS_CONST float *__restrict__ rcp = const_ptr(d.rcp + tb);
    const TVS_CONST float *__restrict__ rcp_lo = const_ptr(d.rcp_lo + tb);
    const TVS_CONST float *__restrict__ curv = const_ptr(d.curv + tb);
    const TVS_CONST float *__restrict__ wt = const_ptr(d.wt + tb);

    // `max_angle = v ? angle : max_angle` is f32::max(angle, max_angle)
    // (kernel.rs:233) for every input the walk meets: v is `angle >
    // max_angle`, so a NaN angle (false) keeps the max as maxNum does.
    float max_angle = kMaxAngle;
    float surface = 0.0f;
    bool visible = true;

    // Ring records (kernel.rs:141-243): an event is written exactly when the
    // visibility changes (opening and closing are exclusive), plus a final
    // close if the last step is visible.  Modes: fixed R-slot records
    // (kRingRecords), per-band event count only (kRingCount), or the events
    // packed at per-band offsets (kRingCompact).
    const bool capture = CAPTURE && active && rel < d.capture_slots;
    const uint32_t kernel_id = backward ? t + d.N : t;
    TVS_GLOBAL uint32_t *band_ring = nullptr;
    uint32_t ring_id = 1;
    unsigned dropped = 0;
    if (CAPTURE && capture) {
        if (d.ring_mode == kRingRecords) {
            band_ring = global_ptr(d.ring) + static_cast<size_t>(rel) * 2u * d.N * d.R +
                        static_cast<size_t>(kernel_id) * d.R;
            if (1u < d.R) band_ring[1] = pov; else ++dropped;
        } else if (d.ring_mode == kRingCompact) {
            band_ring = global_ptr(d.ring_events) + global_ptr(d.ring_offsets)[kernel_id];
        }
    }
    auto event = [&](uint32_t p) {
        ++ring_id;
        // p / 2 is bias + pov_shift +/- off, plus copy_offset for a copy-B
        // step (bias +/- off <= 2 bias < copy_offset; no step goes there
        // when the offset is 0).
        const uint32_t e = p >> 1;
        const uint32_t id = pov - d.pov_shift +
                            (e > 2u * bias + d.pov_shift ? e - d.copy_offset : e) - bias;   // pov +/- off
        if (d.ring_mode == kRingRecords) {
            if (ring_id < d.R) band_ring[ring_id] = id; else ++dropped;
        } else if (d.ring_mode == kRingCompact) {
            band_ring[ring_id - 2u] = id;
        }
    };

    // One step with the IEEE divide (fast_k false) or the sector's divide.
    auto single = [&](uint32_t k, bool fast_k) {
        const uint32_t o = pos[k];
        const float dk = dist[k];
        const float dh = load_height<H>(rsrc, pov_bytes, soff<H>(o)) - h0;
        const float q = (DIV != 0 && fast_k) ? step_div<DIV>(dh, dk, rcp[k], rcp_lo[k]) : dh / dk;
        const float angle = q - curv[k];
        const bool v = angle > max_angle;
        if (CAPTURE && capture && v != visible) event(o);
        surface = add_if(v, surface, wt[k]);
        max_angle = v ? angle : max_angle;
        visible = v;
    };
    uint32_t k = 0;
    for (; k < head; ++k) single(k, false);
    // The fast range [head, n_fast) in whole chunks, through the tables
    // from step head on (f*), indexed from 0.
    const uint32_t fast_end = (n_fast - head) - (n_fast - head) % UNROLL;
    const TVS_CONST uint32_t *__restrict__ fpos = pos + head;
    const TVS_CONST float *__restrict__ fdist = dist + head;
    const TVS_CONST float *__restrict__ frcp = rcp + head;
    const TVS_CONST float *__restrict__ frcp_lo = rcp_lo + head;
    const TVS_CONST float *__restrict__ fcurv = curv + head;
    const TVS_CONST float *__restrict__ fwt = wt + head;
    auto chunk = [&](const float (&h)[UNROLL], uint32_t k0) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const float dh = h[u] - h0;
            const float q = step_div<DIV>(dh, fdist[k0 + u], frcp[k0 + u], frcp_lo[k0 + u]);
            const float angle = q - fcurv[k0 + u];
            if (!CAPTURE) {
                step_update(angle, fwt[k0 + u], lanes, max_angle, surface);
                continue;
            }
            const bool v = angle > max_angle;
            if (capture && v != visible) event(fpos[k0 + u]);
            surface = add_if(v, surface, fwt[k0 + u]);
            max_angle = v ? angle : max_angle;
            visible = v;
        }
    };
    uint32_t j = 0;
    if (PIPE) {
        // Heights of the next chunk are requested before this chunk's
        // arithmetic; the last iteration re-reads its own chunk (in range).
        if (fast_end) {
            float h[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) h[u] = load_height<H>(rsrc, pov_bytes, soff<H>(fpos[u]));
            for (; j < fast_end; j += UNROLL) {
                const uint32_t jn = j + UNROLL < fast_end ? j + UNROLL : j;
                float hn[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u)
                    hn[u] = load_height<H>(rsrc, pov_bytes, soff<H>(fpos[jn + u]));
                chunk(h, j);
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) h[u] = hn[u];
            }
        }
    } else {
        for (; j < fast_end; j += UNROLL) {
            float h[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) h[u] = load_height<H>(rsrc, pov_bytes, soff<H>(fpos[j + u]));
            chunk(h, j);
        }
    }
    k += fast_end;
    for (; k < tail_start; ++k) single(k, k < n_fast);
    if (k < d.n_steps) {
        // Zero tail: delta 0 keeps the point, distance +0 gives the same
        // angle (dh/0 - 0: +inf, -inf or NaN) at every step, and the weight
        // 0 * 0.0174533 leaves the surface unchanged.  Only its first step
        // can see the point: after it the max is at least the angle (or the
        // angle is NaN), so every later step is hidden -- one event at most
        // at the second step, none after.
        const uint32_t o = pos[k];
        const float dh = load_height<H>(rsrc, pov_bytes, soff<H>(o)) - h0;
        const float angle = dh / dist[k] - curv[k];
        const bool v = angle > max_angle;
        if (CAPTURE && capture && v != visible) event(o);
        max_angle = v ? angle : max_angle;
        visible = v;
        if (k + 1u < d.n_steps) {
            if (CAPTURE && capture && visible) event(o);
            visible = false;
        }
    }
    if (CAPTURE && capture) {
        if (visible) event(d.n_steps ? pos[d.n_steps - 1] : 2u * (bias + d.pov_shift));
        if (d.ring_mode == kRingRecords) {
            if (d.R > 0) band_ring[0] = ring_id;
            if (dropped) atomicAdd(d.overflow, static_cast<unsigned long long>(dropped));
        } else if (d.ring_mode == kRingCount) {
            global_ptr(d.ring_counts)[kernel_id] = ring_id;
        }
    }
    const uint32_t pass = 2u * rel + (backward ? 1u : 0u);
    if (active) global_ptr(d.partial)[static_cast<size_t>(pass) * d.N + t] = surface;
}

// Both bands of point t in sector slot d.first_slot + rel in one lane, no
// ring capture.  The forward and backward walks share every per-step
// constant (distance, reciprocal, curvature, weight) and differ only in the
// step's position, so each scalar table read and each loop iteration's
// bookkeeping serves two bands, and the lane carries two independent
// max-scan chains.  The zero tail is skipped: its angle is the same at every
// step and its weight is +0, so it changes neither surface.
template <typename H, int DIV, int UNROLL>
__device__ __forceinline__ void walk_pair(const TileDesc &d, uint32_t rel, uint32_t t,
                                          bool active) {
    const H *elev = static_cast<const H *>(d.elev);
    const uint32_t slot = d.first_slot + rel;
    const uint32_t pov = (t / d.T + d.L) * d.w + (t % d.T + d.L);
    const uint32_t pov_bytes = (pov - d.pov_shift) * static_cast<uint32_t>(sizeof(H));
    const float h0 = static_cast<float>(global_ptr(elev)[pov]) + d.obs;
    const uint32_t head = const_ptr(d.meta)[slot].head;
    const uint32_t n_fast = const_ptr(d.meta)[slot].n_fast;
    const uint32_t tail_start = const_ptr(d.meta)[slot].tail_start;
    const uint32_t bias = const_ptr(d.meta)[slot].bias;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<H *>(elev - bias), 0,
        static_cast<int>((bias + d.span) * sizeof(H)), kHeightRsrcFlags<H>);

    const size_t tb = static_cast<size_t>(slot) * d.stride;
    const TVS_CONST uint32_t *__restrict__ pos_f = const_ptr(d.pos_f + tb);
    const TVS_CONST uint32_t *__restrict__ pos_b = const_ptr(d.pos_b + tb);
    const uint64_t lanes = __builtin_amdgcn_read_exec();
    const TVS_CONST float *__restrict__ dist = const_ptr(d.dist + tb);
    const TVS_CONST float *__restrict__ rcp = const_ptr(d.rcp + tb);
    const TVS_CONST float *__restrict__ rcp_lo = const_ptr(d.rcp_lo + tb);
    const TVS_CONST float *__restrict__ curv = const_ptr(d.curv + tb);
    const TVS_CONST float *__restrict__ wt = const_ptr(d.wt + tb);

    float max_f = kMaxAngle, max_b = kMaxAngle;
    float surface_f = 0.0f, surface_b = 0.0f;
    // One step with the IEEE divide (fast_k false) or the sector's divide.
    auto single = [&](uint32_t k, bool fast_k) {
        const float dk = dist[k];
        const float dhf = load_height<H>(rsrc, pov_bytes, soff<H>(pos_f[k])) - h0;
        const float dhb = load_height<H>(rsrc, pov_bytes, soff<H>(pos_b[k])) - h0;
        const float qf = (DIV != 0 && fast_k) ? step_div<DIV>(dhf, dk, rcp[k], rcp_lo[k]) : dhf / dk;
        const float qb = (DIV != 0 && fast_k) ? step_div<DIV>(dhb, dk, rcp[k], rcp_lo[k]) : dhb / dk;
        pair_update(qf - curv[k], qb - curv[k], wt[k], lanes, max_f, max_b, surface_f,
                    surface_b);
    };
    uint32_t k = 0;
    for (; k < head; ++k) single(k, false);
    // The fast range [head, n_fast) in whole chunks, through the tables
    // from step head on (f*), indexed from 0.
    const uint32_t fast_end = (n_fast - head) - (n_fast - head) % UNROLL;
    const TVS_CONST uint32_t *__restrict__ fpos_f = pos_f + head;
    const TVS_CONST uint32_t *__restrict__ fpos_b = pos_b + head;
    const TVS_CONST float *__restrict__ fdist = dist + head;
    const TVS_CONST float *__restrict__ frcp = rcp + head;
    const TVS_CONST float *__restrict__ frcp_lo = rcp_lo + head;
    const TVS_CONST float *__restrict__ fcurv = curv + head;
    const TVS_CONST float *__restrict__ fwt = wt + head;
    for (uint32_t j = 0; j < fast_end; j += UNROLL) {
        float hf[UNROLL], hb[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            hf[u] = load_height<H>(rsrc, pov_bytes, soff<H>(fpos_f[j + u]));
            hb[u] = load_height<H>(rsrc, pov_bytes, soff<H>(fpos_b[j + u]));
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const float dhf = hf[u] - h0;
            const float dhb = hb[u] - h0;
            const float qf = step_div<DIV>(dhf, fdist[j + u], frcp[j + u], frcp_lo[j + u]);
            const float qb = step_div<DIV>(dhb, fdist[j + u], frcp[j + u], frcp_lo[j + u]);
            pair_update(qf - fcurv[j + u], qb - fcurv[j + u], fwt[j + u], lanes, max_f, max_b,
                        surface_f, surface_b);
        }
    }
    k += fast_end;
    for (; k < tail_start; ++k) single(k, k < n_fast);
    if (active) {
        TVS_GLOBAL float *out = global_ptr(d.partial) + static_cast<size_t>(2u * rel) * d.N + t;
        out[0] = surface_f;
        out[d.N] = surface_b;
    }
}

// Heights of two adjacent elements (the step's position for a lane's point
// t0 and for t0 + 1, which sit next to each other in the row).  int16: one
// dword load returns both, so a band-step costs half a load instruction
// (the texture unit's rate per load instruction bounds this walk); the
// step tables point the steps whose pair would start at an odd element
// into the DEM's second copy, which starts at an odd element, so every such
// dword is 4-byte aligned (a dword at 2 mod 4 costs the unit twice).  f32:
// two loads.  `raw` is an untyped resource over the same range as the
// heights' typed one.
template <typename H>
__device__ __forceinline__ void load_pair(__amdgpu_buffer_rsrc_t raw, uint32_t pov_bytes,
                                          uint32_t pos16, float &a, float &b);

template <>
__device__ __forceinline__ void load_pair<int16_t>(__amdgpu_buffer_rsrc_t raw, uint32_t pov_bytes,
                                                   uint32_t pos16, float &a, float &b) {
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(raw, pov_bytes, pos16, 0);
    a = static_cast<float>(static_cast<int16_t>(w & 0xffffu));
    b = static_cast<float>(static_cast<int32_t>(w) >> 16);
}

template <>
__device__ __forceinline__ void load_pair<float>(__amdgpu_buffer_rsrc_t raw, uint32_t pov_bytes,
                                                 uint32_t pos16, float &a, float &b) {
    a = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(raw, pov_bytes, pos16 << 1, 0));
    b = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(raw, pov_bytes + 4u, pos16 << 1, 0));
}

// Two adjacent points of a row per lane, t0 and t0 + 1, in one direction
// or (BOTH) in both, no ring capture.  Bands per lane: point t0 and t0 + 1
// forward-or-`backward`, then (BOTH) t0 and t0 + 1 backward.  Every band
// keeps its own max and surface; the per-step constants and table reads
// are shared by all of them.  The zero tail is skipped as in walk_pair.
template <typename H, int DIV, int UNROLL, bool BOTH>
__device__ __forceinline__ void walk_twin(const TileDesc &d, uint32_t rel, bool backward,
                                          uint32_t t0, bool active) {
    constexpr int NB = BOTH ? 4 : 2;
    const H *elev = static_cast<const H *>(d.elev);
    const uint32_t slot = d.first_slot + rel;
    const uint32_t pov = (t0 / d.T + d.L) * d.w + (t0 % d.T + d.L);
    const uint32_t pov_bytes = (pov - d.pov_shift) * static_cast<uint32_t>(sizeof(H));
    const float h0a = static_cast<float>(global_ptr(elev)[pov]) + d.obs;
    const float h0b = static_cast<float>(global_ptr(elev)[pov + 1u]) + d.obs;
    const uint32_t head = const_ptr(d.meta)[slot].head;
    const uint32_t n_fast = const_ptr(d.meta)[slot].n_fast;
    const uint32_t tail_start = const_ptr(d.meta)[slot].tail_start;
    const uint32_t bias = const_ptr(d.meta)[slot].bias;
    const __amdgpu_buffer_rsrc_t raw = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<H *>(elev - bias), 0,
        static_cast<int>((bias + d.span) * sizeof(H)), kHeightRsrcFlags<float>);

    const size_t tb = static_cast<size_t>(slot) * d.stride;
    const TVS_CONST uint32_t *__restrict__ pos0 =
        const_ptr(((BOTH || !backward) ? d.pos_f : d.pos_b) + tb);
    const TVS_CONST uint32_t *__restrict__ pos1 = const_ptr(d.pos_b + tb);
    const TVS_CONST float *__restrict__ dist = const_ptr(d.dist + tb);
    const TVS_CONST float *__restrict__ rcp = const_ptr(d.rcp + tb);
    const TVS_CONST float *__restrict__ rcp_lo = const_ptr(d.rcp_lo + tb);
    const TVS_CONST float *__restrict__ curv = const_ptr(d.curv + tb);
    const TVS_CONST float *__restrict__ wt = const_ptr(d.wt + tb);
    const uint64_t lanes = __builtin_amdgcn_read_exec();

    float max_angle[NB], surface[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        max_angle[b] = kMaxAngle;
        surface[b] = 0.0f;
    }
    // One step with the IEEE divide (fast_k false) or the sector's divide.
    auto single = [&](uint32_t k, bool fast_k) {
        float h[NB];
        load_pair<H>(raw, pov_bytes, pos0[k], h[0], h[1]);
        if (BOTH) load_pair<H>(raw, pov_bytes, pos1[k], h[NB - 2], h[NB - 1]);
        const float dk = dist[k];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const float dh = h[b] - ((b & 1) ? h0b : h0a);
            const float q = (DIV != 0 && fast_k) ? step_div<DIV>(dh, dk, rcp[k], rcp_lo[k]) : dh / dk;
            step_update(q - curv[k], wt[k], lanes, max_angle[b], surface[b]);
        }
    };
    uint32_t k = 0;
    for (; k < head; ++k) single(k, false);
    // The fast range [head, n_fast) in whole chunks, through the tables
    // from step head on (f*), indexed from 0.
    const uint32_t fast_end = (n_fast - head) - (n_fast - head) % UNROLL;
    const TVS_CONST uint32_t *__restrict__ fpos0 = pos0 + head;
    const TVS_CONST uint32_t *__restrict__ fpos1 = pos1 + head;
    const TVS_CONST float *__restrict__ fdist = dist + head;
    const TVS_CONST float *__restrict__ frcp = rcp + head;
    const TVS_CONST float *__restrict__ frcp_lo = rcp_lo + head;
    const TVS_CONST float *__restrict__ fcurv = curv + head;
    const TVS_CONST float *__restrict__ fwt = wt + head;
    // int16 heights with their f16 form, both directions: each height
    // difference is one mixed-precision FMA from the f16 dword (the same
    // bytes and load as the int16 one) when every observer of the wave has
    // an exact C - h0 (about 98 % of the waves at 1201 x 1201, heights
    // [0, 3000], C = 952); else the int16 heights are converted.  The steps
    // outside the unrolled range read the int16 heights.
    bool half = false;
    if constexpr (sizeof(H) == 2 && BOTH && DIV != 0) {
        float ca = 0.0f, cb = 0.0f;
        half = d.elev_h16 != nullptr;
        if (half) {
            const bool ok_a = exact_offset(d.h16_offset, h0a, ca);
            const bool ok_b = exact_offset(d.h16_offset, h0b, cb);
            half = __builtin_amdgcn_ballot_w64(!(ok_a && ok_b)) == 0;
        }
        if (half) {
            const __amdgpu_buffer_rsrc_t hraw = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<void *>(static_cast<const void *>(
                    static_cast<const uint16_t *>(d.elev_h16) - bias)),
                0, static_cast<int>((bias + d.span) * sizeof(H)), kHeightRsrcFlags<float>);
            for (uint32_t j = 0; j < fast_end; j += UNROLL) {
                uint32_t hw[UNROLL][2];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    hw[u][0] = __builtin_amdgcn_raw_buffer_load_b32(hraw, pov_bytes, fpos0[j + u], 0);
                    hw[u][1] = __builtin_amdgcn_raw_buffer_load_b32(hraw, pov_bytes, fpos1[j + u], 0);
                }
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) {
                    // With the conversion gone, copying the reciprocal's
                    // high part into a vector register as well as the
                    // curvature term measured ~1 % faster (both parts: no
                    // gain; neither: 3 % slower).
                    const float dk = fdist[j + u];
                    const float y_hi = to_vgpr(frcp[j + u]);
                    const float y_lo = frcp_lo[j + u];
                    const float c = to_vgpr(fcurv[j + u]);
                    const float w = fwt[j + u];
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const uint32_t x = hw[u][b >> 1];
                        const float dh = (b & 1) ? half_hi_diff(x, cb) : half_lo_diff(x, ca);
                        step_update(step_div<DIV>(dh, dk, y_hi, y_lo) - c, w, lanes,
                                    max_angle[b], surface[b]);
                    }
                }
            }
        }
    }
    for (uint32_t j = 0; !half && j < fast_end; j += UNROLL) {
        float h[UNROLL][NB];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            load_pair<H>(raw, pov_bytes, fpos0[j + u], h[u][0], h[u][1]);
            if (BOTH) load_pair<H>(raw, pov_bytes, fpos1[j + u], h[u][NB - 2], h[u][NB - 1]);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const float dk = fdist[j + u];
            const float y_hi = frcp[j + u];
            const float y_lo = frcp_lo[j + u];
            // Two vector instructions issue in one quad-cycle only from two
            // waves, and never when both read a scalar register: with four
            // bands per lane the curvature term is copied into a vector
            // register once per step (a quarter instruction per band-step),
            // which leaves three of the step's eight instructions with a
            // scalar operand instead of four.  Measured 1-1.5 % faster on
            // 1201 x 1201 (same box); copying the reciprocal pair or the
            // weight as well, or instead, gained less or lost.
            const float c = BOTH ? to_vgpr(fcurv[j + u]) : fcurv[j + u];
            const float w = fwt[j + u];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float dh = h[u][b] - ((b & 1) ? h0b : h0a);
                step_update(step_div<DIV>(dh, dk, y_hi, y_lo) - c, w, lanes, max_angle[b],
                            surface[b]);
            }
        }
    }
    k += fast_end;
    for (; k < tail_start; ++k) single(k, k < n_fast);
    if (active) {
        TVS_GLOBAL float *out =
            global_ptr(d.partial) + static_cast<size_t>(2u * rel + (BOTH ? 0u : (backward ? 1u : 0u))) * d.N + t0;
        out[0] = surface[0];
        out[1] = surface[1];
        if (BOTH) {
            out[d.N] = surface[NB - 2];
            out[d.N + 1] = surface[NB - 1];
        }
    }
}

// Two pairs of adjacent points per lane, (ta, ta + 1) of one row and (tb,
// tb + 1) of the next, in both directions: eight bands share every step's
// constants, which are copied into vector registers once per step (four
// moves for eight bands) so that the step's arithmetic reads no scalar
// register.  Bands: a, a + 1, b, b + 1 forward, then the same backward.
// `active_b` false: the lane's second pair duplicates the first (no row
// below it) and is not written.  Regular steps only use the fast divides
// (DIV 2 or 4); the zero tail is skipped as in walk_pair.
template <typename H, int DIV, int UNROLL>
__device__ __forceinline__ void walk_quad(const TileDesc &d, uint32_t rel, uint32_t ta,
                                          uint32_t tb, bool active_a, bool active_b) {
    static_assert(DIV == 2 || DIV == 4, "walk_quad takes the fast divides");
    constexpr int NB = 8;
    const H *elev = static_cast<const H *>(d.elev);
    const uint32_t slot = d.first_slot + rel;
    const uint32_t pov_a = (ta / d.T + d.L) * d.w + (ta % d.T + d.L);
    const uint32_t pov_b = (tb / d.T + d.L) * d.w + (tb % d.T + d.L);
    const uint32_t bytes_a = (pov_a - d.pov_shift) * static_cast<uint32_t>(sizeof(H));
    const uint32_t bytes_b = (pov_b - d.pov_shift) * static_cast<uint32_t>(sizeof(H));
    float h0[4];
    h0[0] = static_cast<float>(global_ptr(elev)[pov_a]) + d.obs;
    h0[1] = static_cast<float>(global_ptr(elev)[pov_a + 1u]) + d.obs;
    h0[2] = static_cast<float>(global_ptr(elev)[pov_b]) + d.obs;
    h0[3] = static_cast<float>(global_ptr(elev)[pov_b + 1u]) + d.obs;
    const uint32_t head = const_ptr(d.meta)[slot].head;
    const uint32_t n_fast = const_ptr(d.meta)[slot].n_fast;
    const uint32_t tail_start = const_ptr(d.meta)[slot].tail_start;
    const uint32_t bias = const_ptr(d.meta)[slot].bias;
    const __amdgpu_buffer_rsrc_t raw = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<H *>(elev - bias), 0,
        static_cast<int>((bias + d.span) * sizeof(H)), kHeightRsrcFlags<float>);

    const size_t tb_ = static_cast<size_t>(slot) * d.stride;
    const TVS_CONST uint32_t *__restrict__ pos_f = const_ptr(d.pos_f + tb_);
    const TVS_CONST uint32_t *__restrict__ pos_b = const_ptr(d.pos_b + tb_);
    const TVS_CONST float *__restrict__ dist = const_ptr(d.dist + tb_);
    const TVS_CONST float *__restrict__ rcp = const_ptr(d.rcp + tb_);
    const TVS_CONST float *__restrict__ rcp_lo = const_ptr(d.rcp_lo + tb_);
    const TVS_CONST float *__restrict__ curv = const_ptr(d.curv + tb_);
    const TVS_CONST float *__restrict__ wt = const_ptr(d.wt + tb_);
    const uint64_t lanes = __builtin_amdgcn_read_exec();

    float max_angle[NB], surface[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        max_angle[b] = kMaxAngle;
        surface[b] = 0.0f;
    }
    auto load8 = [&](uint32_t pf, uint32_t pb, float (&h)[NB]) {
        load_pair<H>(raw, bytes_a, pf, h[0], h[1]);
        load_pair<H>(raw, bytes_b, pf, h[2], h[3]);
        load_pair<H>(raw, bytes_a, pb, h[4], h[5]);
        load_pair<H>(raw, bytes_b, pb, h[6], h[7]);
    };
    // One step with the IEEE divide (fast_k false) or the sector's divide.
    auto single = [&](uint32_t k, bool fast_k) {
        float h[NB];
        load8(pos_f[k], pos_b[k], h);
        const float dk = dist[k];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const float dh = h[b] - h0[b & 3];
            const float q = fast_k ? step_div<DIV>(dh, dk, rcp[k], rcp_lo[k]) : dh / dk;
            step_update(q - curv[k], wt[k], lanes, max_angle[b], surface[b]);
        }
    };
    uint32_t k = 0;
    for (; k < head; ++k) single(k, false);
    const uint32_t fast_end = (n_fast - head) - (n_fast - head) % UNROLL;
    const TVS_CONST uint32_t *__restrict__ fpos_f = pos_f + head;
    const TVS_CONST uint32_t *__restrict__ fpos_b = pos_b + head;
    const TVS_CONST float *__restrict__ fdist = dist + head;
    const TVS_CONST float *__restrict__ frcp = rcp + head;
    const TVS_CONST float *__restrict__ frcp_lo = rcp_lo + head;
    const TVS_CONST float *__restrict__ fcurv = curv + head;
    const TVS_CONST float *__restrict__ fwt = wt + head;
    for (uint32_t j = 0; j < fast_end; j += UNROLL) {
        float h[UNROLL][NB];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) load8(fpos_f[j + u], fpos_b[j + u], h[u]);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const float dk = DIV == 4 ? to_vgpr(fdist[j + u]) : 0.0f;
            const float y_hi = to_vgpr(frcp[j + u]);
            const float y_lo = to_vgpr(frcp_lo[j + u]);
            const float c = to_vgpr(fcurv[j + u]);
            const float w = to_vgpr(fwt[j + u]);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float dh = h[u][b] - h0[b & 3];
                step_update_v(step_div<DIV>(dh, dk, y_hi, y_lo) - c, w, lanes, max_angle[b],
                              surface[b]);
            }
        }
    }
    k += fast_end;
    for (; k < tail_start; ++k) single(k, k < n_fast);
    TVS_GLOBAL float *out = global_ptr(d.partial) + static_cast<size_t>(2u * rel) * d.N;
    if (active_a) {
        out[ta] = surface[0];
        out[ta + 1u] = surface[1];
        out[d.N + ta] = surface[4];
        out[d.N + ta + 1u] = surface[5];
    }
    if (active_b) {
        out[tb] = surface[2];
        out[tb + 1u] = surface[3];
        out[d.N + tb] = surface[6];
        out[d.N + tb + 1u] = surface[7];
    }
}

// Index of the tile whose [begin[i], begin[i+1]) holds x (wave-uniform).
__device__ __forceinline__ uint32_t find_tile(const uint32_t *__restrict__ begin,
                                              uint32_t n_tiles, uint32_t x) {
    uint32_t lo = 0, hi = n_tiles;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (begin[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Divide of a sector's regular steps (wave-uniform): the short divide where
// it was verified for this DEM, else the split-reciprocal divide where the
// heights passed the range check, else IEEE.
__device__ __forceinline__ int divide_of(const TileDesc &d, uint32_t rel) {
    if (!d.fast) return 0;
    if (d.short_div_bad && const_ptr(d.short_div_bad)[d.first_slot + rel] == 0u) return 2;
    return 4;
}

// Divide variants of one launch.  IEEE: every tile walks with the IEEE
// divide (no tile passed the fast-divide checks), at the launch's unroll.
// Otherwise the fast divides run at the launch's unroll and a tile without
// them (batches may mix) walks with the IEEE divide one step at a time: the
// unrolled IEEE loop holds so many temporaries that the whole kernel's
// register count, and so the waves per SIMD of the fast loops, would follow
// it (94 against 57 VGPRs for two points and both directions per lane).
//
// A work item is (tile, sector, direction, BLOCK points).  With PAIR an item
// is (tile, sector, BLOCK points), both directions: `item` counts pairs and
// 2*item is the first of the pair's two single-direction items (every tile
// holds an even number of them).  With TWIN the points are taken two per
// lane: the first twin_blocks items of a pass hold the column pairs (2c,
// 2c + 1) of every row (T/2 pairs per row), the rest (odd T only) the
// column left over in each row, one point per lane.  With an odd DEM width
// (odd T) the pairs of odd rows start one column later, (2c + 1, 2c + 2),
// and their row's left-over column is 0 instead of T - 1: then the first
// point of every pair has an even DEM index, so at any step the dword
// loads of a whole wave are either all 4-byte aligned or all at 2 mod 4,
// never mixed (a wave whose lanes span two rows would otherwise pay the
// misaligned rate at every step).
template <typename H, bool CAPTURE, int BLOCK, int UNROLL, bool PIPE, bool PAIR, int TWIN,
          bool IEEE>
__device__ __forceinline__ void walk_item(const TileDesc *__restrict__ tiles,
                                          const uint32_t *__restrict__ item_begin,
                                          uint32_t n_tiles, uint32_t item) {
    const uint32_t x = PAIR ? 2u * item : item;
    const uint32_t ti = n_tiles > 1 ? find_tile(item_begin, n_tiles, x) : 0u;
    const TileDesc d = tiles[ti];
    const uint32_t local = PAIR ? (x - item_begin[ti]) >> 1 : x - item_begin[ti];
    const uint32_t pass = local / d.blocks_per_pass;     // sector (PAIR) or 2*sector + dir
    const uint32_t rel = PAIR ? pass : pass >> 1;
    uint32_t ob = local - pass * d.blocks_per_pass;
    if (d.order)   // directional launch: the sector's blocks in strip order
        ob = const_ptr(d.order)[(d.first_slot + rel) * d.blocks_per_pass + ob];
    const bool backward = !PAIR && (pass & 1u) != 0;
    const int div = IEEE ? 0 : divide_of(d, rel);
    if (TWIN == 2) {
        // Lane groups: the column pairs of rows 2r and 2r + 1 (the last row
        // alone when T is odd), each pair as in TWIN == 1.
        const uint32_t half = d.T >> 1;
        const uint32_t n_groups = twin_lane_groups(d.T, 2u);
        const uint32_t group_blocks = (n_groups + BLOCK - 1) / BLOCK;
        if (ob < group_blocks) {
            const uint32_t q_raw = ob * BLOCK + threadIdx.x;
            const bool active = q_raw < n_groups;
            const uint32_t q = active ? q_raw : n_groups - 1u;
            const uint32_t rp = q / half;
            const uint32_t col = q - rp * half;
            const uint32_t ra = 2u * rp;
            const bool has_b = ra + 1u < d.T;
            const uint32_t rb = has_b ? ra + 1u : ra;
            const uint32_t ta = ra * d.T + 2u * col + (d.w & ra & 1u);
            const uint32_t tb = rb * d.T + 2u * col + (d.w & rb & 1u);
            if (!IEEE && div == 2) {
                walk_quad<H, 2, UNROLL>(d, rel, ta, tb, active, active && has_b);
            } else if (!IEEE && div == 4) {
                walk_quad<H, 4, UNROLL>(d, rel, ta, tb, active, active && has_b);
            } else {
                // IEEE divide: the two pairs one after the other, at one
                // step per iteration (the unrolled IEEE loop of eight bands
                // would hold too many registers).
                walk_twin<H, 0, 1, true>(d, rel, false, ta, active);
                walk_twin<H, 0, 1, true>(d, rel, false, tb, active && has_b);
            }
            return;
        }
    } else if (TWIN) {
        const uint32_t half = d.T >> 1;
        const uint32_t n_pairs = d.T * half;
        const uint32_t twin_blocks = (n_pairs + BLOCK - 1) / BLOCK;
        if (ob < twin_blocks) {
            const uint32_t p_raw = ob * BLOCK + threadIdx.x;
            const bool active = p_raw < n_pairs;
            const uint32_t p = active ? p_raw : n_pairs - 1u;
            const uint32_t row = p / half;
            const uint32_t t0 = row * d.T + 2u * (p - row * half) + (d.w & row & 1u);
            if (IEEE)
                walk_twin<H, 0, UNROLL, PAIR>(d, rel, backward, t0, active);
            else if (div == 2)
                walk_twin<H, 2, UNROLL, PAIR>(d, rel, backward, t0, active);
            else if (div == 4)
                walk_twin<H, 4, UNROLL, PAIR>(d, rel, backward, t0, active);
            else
                walk_twin<H, 0, 1, PAIR>(d, rel, backward, t0, active);
            return;
        }
    }
    // One point per lane: consecutive points, or (TWIN) each row's left-over
    // column.
    const uint32_t lb = TWIN ? ob - (twin_lane_groups(d.T, TWIN) + BLOCK - 1) / BLOCK : ob;
    const uint32_t i_raw = lb * BLOCK + threadIdx.x;
    const uint32_t n_single = TWIN ? d.T : d.N;
    const bool active = i_raw < n_single;
    const uint32_t i = active ? i_raw : n_single - 1u;   // spare lanes shadow
    const uint32_t t = TWIN ? i * d.T + ((d.w & i & 1u) ? 0u : d.T - 1u) : i;
    if (PAIR) {
        if (IEEE)
            walk_pair<H, 0, UNROLL>(d, rel, t, active);
        else if (div == 2)
            walk_pair<H, 2, UNROLL>(d, rel, t, active);
        else if (div == 4)
            walk_pair<H, 4, UNROLL>(d, rel, t, active);
        else
            walk_pair<H, 0, 1>(d, rel, t, active);
        return;
    }
    if (IEEE)
        walk_band<H, 0, CAPTURE, UNROLL, PIPE>(d, rel, backward, t, active);
    else if (div == 2)
        walk_band<H, 2, CAPTURE, UNROLL, PIPE>(d, rel, backward, t, active);
    else if (div == 4)
        walk_band<H, 4, CAPTURE, UNROLL, PIPE>(d, rel, backward, t, active);
    else
        walk_band<H, 0, CAPTURE, 1, PIPE>(d, rel, backward, t, active);
}

// One work item per block, or (PERSIST) a grid of a few blocks per CU
// striding over all items.  Consecutive block ids are dispatched to the 8
// XCDs round-robin.  With items_per_xcd != 0 they are remapped so each XCD
// takes a contiguous run of items: one DEM region per L2.  With pass_chunk
// != 0 (one tile, pass_items items per pass) each pass's items are split in
// 8 contiguous chunks of pass_chunk, one per XCD: every XCD takes an equal
// share of every sector, and (with the directional block order) its share
// is a strip along the sector's band direction.
template <typename H, bool CAPTURE, int BLOCK, int UNROLL, bool PIPE, bool PERSIST, bool PAIR,
          int TWIN, bool IEEE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TWIN == 2 && sizeof(H) == 2 ? kQuadWaves : 1))) void walk_tiles_kernel(
    const TileDesc *__restrict__ tiles, const uint32_t *__restrict__ item_begin,
    uint32_t n_tiles, uint32_t n_items, uint32_t items_per_xcd, uint32_t pass_items,
    uint32_t pass_chunk) {
    const uint32_t b = blockIdx.x;
    if (PERSIST) {
        for (uint32_t item = b; item < n_items; item += gridDim.x)
            walk_item<H, CAPTURE, BLOCK, UNROLL, PIPE, PAIR, TWIN, IEEE>(tiles, item_begin,
                                                                         n_tiles, item);
        return;
    }
    uint32_t item;
    if (pass_chunk) {
        const uint32_t j = b >> 3;
        const uint32_t pass = j / pass_chunk;
        const uint32_t within = (b & 7u) * pass_chunk + (j - pass * pass_chunk);
        if (within >= pass_items) return;
        item = pass * pass_items + within;
    } else {
        item = items_per_xcd ? (b & 7u) * items_per_xcd + (b >> 3) : b;
    }
    if (item >= n_items) return;
    walk_item<H, CAPTURE, BLOCK, UNROLL, PIPE, PAIR, TWIN, IEEE>(tiles, item_begin, n_tiles, item);
}

// Per point, in sector order: cum += forward; cum += backward (the f32
// order of the reference's serial loop), and the same sums in f64.
__global__ __launch_bounds__(kBlock) void reduce_tiles_kernel(
    const TileDesc *__restrict__ tiles, const uint32_t *__restrict__ point_begin,
    uint32_t n_tiles, uint32_t n_points) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n_points) return;
    const uint32_t ti = n_tiles > 1 ? find_tile(point_begin, n_tiles, g) : 0u;
    const TileDesc &d = tiles[ti];
    const uint32_t N = d.N;
    const uint32_t t = g - point_begin[ti];
    const TVS_GLOBAL float *__restrict__ partial = global_ptr(d.partial);
    TVS_GLOBAL float *cum = global_ptr(d.cum);
    TVS_GLOBAL double *cum64 = global_ptr(d.cum64);
    float c = cum[t];
    double s = cum64[t];
    // Loads for kReduceChunk sectors are issued before their adds (the adds
    // stay in sector order): the kernel is bound by HBM latency otherwise,
    // with only a few waves per SIMD at one point per thread.
    constexpr uint32_t kReduceChunk = 16;
    const size_t stride = 2ull * N;
    const TVS_GLOBAL float *__restrict__ p = partial + t;
    uint32_t r = 0;
    for (; r + kReduceChunk <= d.count; r += kReduceChunk) {
        float f[kReduceChunk], b[kReduceChunk];
#pragma unroll
        for (uint32_t u = 0; u < kReduceChunk; ++u) {
            f[u] = p[(r + u) * stride];
            b[u] = p[(r + u) * stride + N];
        }
#pragma unroll
        for (uint32_t u = 0; u < kReduceChunk; ++u) {
            c = c + f[u];
            c = c + b[u];
            s = s + static_cast<double>(f[u]);
            s = s + static_cast<double>(b[u]);
        }
    }
    for (; r < d.count; ++r) {
        const float f = p[r * stride];
        const float b = p[r * stride + N];
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

// The int16 heights as f16 values h - offset (exact: every |h - offset| <=
// 2048, checked by the caller), as raw bits, over both DEM copies.
__global__ void half_heights_kernel(const int16_t *in, uint16_t *out, size_t n, int32_t offset) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const _Float16 h = static_cast<_Float16>(static_cast<float>(in[i] - offset));
        out[i] = __builtin_bit_cast(uint16_t, h);
    }
}

// Short-divide verification: block (slot, 64-step chunk of the regular
// steps); threads stride over the domain of height differences and compare
// the short divide with the split-reciprocal divide at each of the chunk's
// steps.  Any difference marks the slot (every writer stores 1).
constexpr uint32_t kShortDivChunk = 64;

__global__ __launch_bounds__(kBlock) void short_div_check_kernel(StepTables tab, uint32_t first,
                                                            const float *__restrict__ domain,
                                                            uint32_t n_domain,
                                                            uint32_t *__restrict__ bad) {
    const uint32_t slot = first + blockIdx.x;
    const uint32_t n_fast = tab.meta[slot].n_fast;
    const uint32_t k0 = tab.meta[slot].head + blockIdx.y * kShortDivChunk;
    if (k0 >= n_fast) return;   // uniform over the block
    const uint32_t nk = min(n_fast - k0, kShortDivChunk);
    __shared__ float sd[kShortDivChunk], shi[kShortDivChunk], slo[kShortDivChunk];
    const size_t tb = static_cast<size_t>(slot) * tab.stride + k0;
    for (uint32_t k = threadIdx.x; k < nk; k += blockDim.x) {
        sd[k] = tab.dist[tb + k];
        shi[k] = tab.rcp[tb + k];
        slo[k] = tab.rcp_lo[tb + k];
    }
    __syncthreads();
    uint32_t differ = 0;
    for (uint32_t i = threadIdx.x; i < n_domain; i += blockDim.x) {
        const float a = domain[i];
        for (uint32_t k = 0; k < nk; ++k) {
            const float q4 = div_by_reciprocal(a, sd[k], shi[k], slo[k]);
            const float q2 = div_short(a, shi[k], slo[k]);
            differ |= __float_as_uint(q4) ^ __float_as_uint(q2);
        }
    }
    if (differ) bad[slot] = 1u;
}

// Exclusive scan of the compact ring export's per-band event counts
// (counts[k] - 1: the count includes the PoV slot), in three launches: per
// tile of kScanTile bands the sum, then the tiles' exclusive scan in one
// block (the total last), then each tile's bands.  Integer sums: the order
// of the additions does not matter.
constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanItems = 4;
constexpr uint32_t kScanTile = kScanBlock * kScanItems;

// Inclusive scan of one value per thread over the block (Hillis-Steele).
__device__ __forceinline__ uint64_t block_inclusive_scan(uint64_t v, uint64_t *buf) {
    const uint32_t t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (uint32_t o = 1; o < kScanBlock; o <<= 1) {
        const uint64_t add = t >= o ? buf[t - o] : 0u;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    const uint64_t r = buf[t];
    __syncthreads();   // buf is reused by the caller's next scan
    return r;
}

__global__ __launch_bounds__(kScanBlock) void ring_tile_sums_kernel(const uint32_t *counts,
                                                                    uint32_t n,
                                                                    uint64_t *tile_sums) {
    __shared__ uint64_t buf[kScanBlock];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint64_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanItems; ++j)
        if (base + j < n) s += counts[base + j] - 1u;
    const uint64_t incl = block_inclusive_scan(s, buf);
    if (threadIdx.x == kScanBlock - 1u) tile_sums[blockIdx.x] = incl;
}

__global__ __launch_bounds__(kScanBlock) void ring_tile_scan_kernel(uint64_t *tile_sums,
                                                                    uint32_t tiles,
                                                                    uint64_t *total) {
    __shared__ uint64_t buf[kScanBlock];
    uint64_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kScanBlock) {
        const uint32_t i = b + threadIdx.x;
        const uint64_t v = i < tiles ? tile_sums[i] : 0u;
        const uint64_t incl = block_inclusive_scan(v, buf);
        if (i < tiles) tile_sums[i] = carry + incl - v;
        // The chunk's total, from the last thread's inclusive sum.
        if (threadIdx.x == kScanBlock - 1u) buf[0] = incl;
        __syncthreads();
        carry += buf[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void ring_offsets_kernel(const uint32_t *counts,
                                                                  uint32_t n,
                                                                  const uint64_t *tile_offsets,
                                                                  uint64_t *offsets) {
    __shared__ uint64_t buf[kScanBlock];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanItems; ++j) {
        v[j] = base + j < n ? static_cast<uint64_t>(counts[base + j] - 1u) : 0u;
        s += v[j];
    }
    uint64_t run = tile_offsets[blockIdx.x] + block_inclusive_scan(s, buf) - s;
#pragma unroll
    for (uint32_t j = 0; j < kScanItems; ++j) {
        if (base + j < n) offsets[base + j] = run;
        run += v[j];
    }
}

// Blocks of a persistent launch on the current device: enough for eight
// waves per SIMD.
uint32_t persistent_blocks(uint32_t block) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
        n = 256;
    return static_cast<uint32_t>(n) * (32u * 64u / block);
}

// `n_items` counts single-direction items; a PAIR launch has half as many
// (every tile's count is even).
template <typename H, bool CAPTURE, int BLOCK, int UNROLL, bool PIPE, bool PERSIST, bool PAIR,
          int TWIN, bool IEEE>
hipError_t launch_tiles_v(hipStream_t s, const TileDesc *tiles, const uint32_t *item_begin,
                          uint32_t n_tiles, uint32_t n_items, bool xcd_contiguous,
                          uint32_t pass_items) {
    if (PAIR && (n_items & 1u)) return hipErrorInvalidValue;
    const uint32_t items = PAIR ? n_items / 2u : n_items;
    // Split passes: one tile whose items are whole passes.
    const bool split = !PERSIST && xcd_contiguous && n_tiles == 1u && pass_items != 0u &&
                       items % pass_items == 0u;
    const uint32_t chunk = split ? (pass_items + 7u) / 8u : 0u;
    const uint32_t per_xcd = (xcd_contiguous && !PERSIST && !split) ? (items + 7u) / 8u : 0u;
    uint64_t grid = split ? 8ull * chunk * (items / pass_items) : per_xcd ? per_xcd * 8ull : items;
    if (PERSIST) grid = std::min(items, persistent_blocks(BLOCK));
    if (grid > 0xffffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL((walk_tiles_kernel<H, CAPTURE, BLOCK, UNROLL, PIPE, PERSIST, PAIR, TWIN,
                                          IEEE>),
                       dim3(static_cast<uint32_t>(grid)), dim3(BLOCK), 0, s, tiles, item_begin,
                       n_tiles, items, per_xcd, split ? pass_items : 0u, chunk);
    return hipGetLastError();
}

template <typename H, bool CAPTURE, int BLOCK, int UNROLL, bool PIPE, bool PERSIST, bool PAIR,
          int TWIN>
hipError_t launch_tiles_t(hipStream_t s, const TileDesc *tiles, const uint32_t *item_begin,
                          uint32_t n_tiles, uint32_t n_items, bool xcd_contiguous,
                          bool ieee_only, uint32_t pass_items) {
    return ieee_only
               ? launch_tiles_v<H, CAPTURE, BLOCK, UNROLL, PIPE, PERSIST, PAIR, TWIN, true>(
                     s, tiles, item_begin, n_tiles, n_items, xcd_contiguous, pass_items)
               : launch_tiles_v<H, CAPTURE, BLOCK, UNROLL, PIPE, PERSIST, PAIR, TWIN, false>(
                     s, tiles, item_begin, n_tiles, n_items, xcd_contiguous, pass_items);
}

template <typename H>
hipError_t launch_tiles_h(hipStream_t s, const TileDesc *tiles, const uint32_t *item_begin,
                          uint32_t n_tiles, uint32_t n_items, bool capture, bool ieee_only,
                          const WalkConfig &cfg, uint32_t pass_items) {
    if (capture)   // ring capture is a checking path: one shape (capture_walk_config)
        return launch_tiles_t<H, true, 256, 8, false, false, false, false>(
            s, tiles, item_begin, n_tiles, n_items, cfg.xcd_contiguous, ieee_only, 0u);
    const uint32_t split_items = cfg.directional ? pass_items : 0u;
#define TVS_SHAPE(B, U, P, S, D, W)                                                        \
    if (cfg.block == B && cfg.unroll == U && cfg.pipeline == P && cfg.persistent == S &&   \
        cfg.pair == D && cfg.twin == W)                                                    \
        return launch_tiles_t<H, false, B, U, P, S, D, W>(s, tiles, item_begin, n_tiles,   \
                                                          n_items, cfg.xcd_contiguous,     \
                                                          ieee_only, split_items);
    TVS_WALK_SHAPES(TVS_SHAPE)
#undef TVS_SHAPE
    return hipErrorInvalidValue;
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

bool walk_shape_supported(const WalkConfig &cfg) {
#define TVS_SHAPE(B, U, P, S, D, W)                                                      \
    if (cfg.block == B && cfg.unroll == U && cfg.pipeline == P && cfg.persistent == S &&   \
        cfg.pair == D && cfg.twin == W)                                                    \
        return true;
    TVS_WALK_SHAPES(TVS_SHAPE)
#undef TVS_SHAPE
    return false;
}

hipError_t launch_walk_tiles(hipStream_t s, const TileDesc *tiles, const uint32_t *item_begin,
                             uint32_t n_tiles, uint32_t n_items, bool i16, bool capture,
                             bool ieee_only, const WalkConfig &cfg, uint32_t pass_items) {
    if (n_items == 0) return hipSuccess;
    return i16 ? launch_tiles_h<int16_t>(s, tiles, item_begin, n_tiles, n_items, capture,
                                         ieee_only, cfg, pass_items)
               : launch_tiles_h<float>(s, tiles, item_begin, n_tiles, n_items, capture,
                                       ieee_only, cfg, pass_items);
}

hipError_t launch_short_div_check(hipStream_t s, const StepTables &tab, uint32_t first,
                             uint32_t count, const float *domain, uint32_t n_domain,
                             uint32_t *bad) {
    if (count == 0 || n_domain == 0) return hipSuccess;
    const uint32_t chunks = (tab.stride + kShortDivChunk - 1) / kShortDivChunk;
    if (chunks == 0) return hipSuccess;
    hipLaunchKernelGGL(short_div_check_kernel, dim3(count, chunks), dim3(kBlock), 0, s, tab, first,
                       domain, n_domain, bad);
    return hipGetLastError();
}

hipError_t launch_reduce_tiles(hipStream_t s, const TileDesc *tiles,
                               const uint32_t *point_begin, uint32_t n_tiles,
                               uint32_t n_points) {
    if (n_points == 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_tiles_kernel, dim3((n_points + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, tiles, point_begin, n_tiles, n_points);
    return hipGetLastError();
}

size_t ring_offsets_len(uint32_t n) {
    return static_cast<size_t>(n) + 1u + (n + kScanTile - 1u) / kScanTile;
}

hipError_t launch_ring_offsets(hipStream_t s, const uint32_t *counts, uint32_t n,
                               uint64_t *offsets) {
    if (n == 0) return hipSuccess;
    const uint32_t tiles = (n + kScanTile - 1u) / kScanTile;
    uint64_t *tile_sums = offsets + n + 1u;
    hipLaunchKernelGGL(ring_tile_sums_kernel, dim3(tiles), dim3(kScanBlock), 0, s, counts, n,
                       tile_sums);
    hipLaunchKernelGGL(ring_tile_scan_kernel, dim3(1), dim3(kScanBlock), 0, s, tile_sums, tiles,
                       offsets + n);
    hipLaunchKernelGGL(ring_offsets_kernel, dim3(tiles), dim3(kScanBlock), 0, s, counts, n,
                       tile_sums, offsets);
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

hipError_t launch_half_heights(hipStream_t s, const int16_t *in, uint16_t *out, size_t n,
                               int32_t offset) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(half_heights_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                       0, s, in, out, n, offset);
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
