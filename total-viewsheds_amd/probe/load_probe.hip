// load_probe -- the texture unit's rate for the walk's kinds of height load.
//
// Every CU runs eight waves per SIMD; each wave issues, in a loop, eight
// independent buffer loads per iteration with a wave-uniform scalar offset
// and 64 consecutive int16 elements per wave (the walk's access pattern),
// from a buffer that stays in cache.  Reported: nanoseconds per wave-load
// instruction per CU for a 16-bit typed load (the one-point walk's load), a
// 16-bit raw load, a 4-byte-aligned dword (two heights, the two-point walk's
// load when aligned), a dword at 2 mod 4 (misaligned), a two-component
// typed load, and a dwordx2 (two f32 or four int16 heights) at each
// 2-byte alignment.
//
// Build: make -C total-viewsheds_amd probe.  Run: total-viewsheds_amd/load_probe
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

__device__ float tvs_probe_load_format_x(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                         int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.format.f32");
typedef float tvs_probe_float2 __attribute__((ext_vector_type(2)));
__device__ tvs_probe_float2 tvs_probe_load_format_xy(__amdgpu_buffer_rsrc_t rsrc, int voffset,
                                                      int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.format.v2f32");

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kSpanBytes = 1u << 21;   // scalar offsets stay below this
constexpr int kFormat16 = 4 | (3 << 12) | (2 << 15);             // X, SSCALED, 16
constexpr int kFormat16x2 = 4 | (5 << 3) | (3 << 12) | (5 << 15); // XY, SSCALED, 16_16
constexpr int kRaw = 0x00020000;

// mode 0: format_x 16-bit; 1: raw 16-bit; 2: aligned dword; 3: dword at
// 2 mod 4; 4: format_xy 16_16.
template <int MODE>
__global__ __launch_bounds__(kBlock) void load_kernel(const int16_t *buf, float *out, int iters) {
    const int flags = MODE == 0 ? kFormat16 : MODE == 4 ? kFormat16x2 : kRaw;
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int16_t *>(buf), 0, static_cast<int>(kSpanBytes + 1024), flags);
    const uint32_t lane = threadIdx.x & 63u;
    float acc = 0.0f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t so = ((static_cast<uint32_t>(i * 8 + u) * 40503u) & (kSpanBytes - 1u)) & ~3u;
            if (MODE == 0) acc += tvs_probe_load_format_x(r, lane * 2, so, 0);
            if (MODE == 1)
                acc += static_cast<float>(
                    static_cast<int16_t>(__builtin_amdgcn_raw_buffer_load_b16(r, lane * 2, so, 0)));
            if (MODE == 2)
                acc += __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, lane * 4, so, 0) &
                                       0x3fffffffu);
            if (MODE == 3)
                acc += __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, lane * 4, so + 2, 0) &
                                       0x3fffffffu);
            if (MODE == 4) {
                const tvs_probe_float2 t = tvs_probe_load_format_xy(r, lane * 4, so, 0);
                acc += t.x + t.y;
            }
            if (MODE >= 5) {   // dwordx2: 8-aligned, 4 mod 8, 2 mod 8, 6 mod 8
                const uint32_t mis = MODE == 6 ? 4u : MODE == 7 ? 2u : MODE == 8 ? 6u : 0u;
                const auto t = __builtin_amdgcn_raw_buffer_load_b64(r, lane * 8, so + mis, 0);
                acc += __uint_as_float((t[0] ^ t[1]) & 0x3fffffffu);
            }
        }
    }
    out[blockIdx.x * kBlock + threadIdx.x] = acc;
}

template <int MODE>
int run(const int16_t *buf, float *out, int blocks, int cus, int iters, const char *name) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    hipLaunchKernelGGL(load_kernel<MODE>, dim3(blocks), dim3(kBlock), 0, 0, buf, out, iters);
    if (hipEventRecord(e0) != hipSuccess) return 1;
    hipLaunchKernelGGL(load_kernel<MODE>, dim3(blocks), dim3(kBlock), 0, 0, buf, out, iters);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    const double loads_per_cu = blocks * (kBlock / 64.0) * iters * 8.0 / cus;
    std::printf("{\"probe\": \"%s\", \"ms\": %.4f, \"ns_per_wave_load_per_cu\": %.4f}\n", name,
                ms, ms * 1e6 / loads_per_cu);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // namespace

int main() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    const int blocks = cus * 8;
    int16_t *buf = nullptr;
    float *out = nullptr;
    if (hipMalloc(&buf, kSpanBytes + 4096) != hipSuccess) return 1;
    if (hipMemset(buf, 1, kSpanBytes + 4096) != hipSuccess) return 1;
    if (hipMalloc(&out, static_cast<size_t>(blocks) * kBlock * sizeof(float)) != hipSuccess)
        return 1;
    int rc = run<0>(buf, out, blocks, cus, 2000, "16-bit typed (format_x, SSCALED)");
    rc |= run<1>(buf, out, blocks, cus, 2000, "16-bit raw");
    rc |= run<2>(buf, out, blocks, cus, 2000, "dword, 4-byte aligned (two heights)");
    rc |= run<3>(buf, out, blocks, cus, 2000, "dword at 2 mod 4 (two heights, misaligned)");
    rc |= run<4>(buf, out, blocks, cus, 2000, "typed 16_16 (format_xy, two heights)");
    rc |= run<5>(buf, out, blocks, cus, 2000, "dwordx2, 8-byte aligned (two f32 heights)");
    rc |= run<6>(buf, out, blocks, cus, 2000, "dwordx2 at 4 mod 8 (two f32 heights)");
    rc |= run<7>(buf, out, blocks, cus, 2000, "dwordx2 at 2 mod 8 (four int16 heights)");
    rc |= run<8>(buf, out, blocks, cus, 2000, "dwordx2 at 6 mod 8 (four int16 heights)");
    (void)hipFree(buf);
    (void)hipFree(out);
    return rc;
}
