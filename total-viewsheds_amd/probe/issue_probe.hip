// issue_probe -- the instruction-issue ceiling of the batched walk's step.
//
// Runs the walk's per-step vector sequence (sub, mul, fmac, fma, fmac,
// subrev, v_cmpx, masked add, masked move, exec restore: 9 VALU + 1 SALU;
// 7 VALU with the short divide) out of registers, with no memory access, on every SIMD at full occupancy, and
// reports nanoseconds per wave-instruction per SIMD.  Next to it: the same
// count of independent FMAs and of independent adds.  The walk kernel's own
// rate (kernel time / its VALU instructions per SIMD) is compared with these
// in the measurement notes; the gap is what the height loads and the loop's
// scalar work cost.
//
// Build: make -C total-viewsheds_amd probe.  Run: total-viewsheds_amd/issue_probe
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlock = 256;

#define TVS_REP8(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)

// mode 0: the walk step; 1: independent v_fma_f32; 2: independent v_add_f32;
// 3: the walk step with the short divide.
template <int MODE>
__global__ __launch_bounds__(kBlock) void issue_kernel(float *out, int iters, float c0,
                                                       float c1) {
    float x0 = threadIdx.x * 1e-3f, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4,
          x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
    float y0 = x0 * 0.5f, y1 = x1 * 0.5f, y2 = x2 * 0.5f, y3 = x3 * 0.5f, y4 = x4 * 0.5f,
          y5 = x5 * 0.5f, y6 = x6 * 0.5f, y7 = x7 * 0.5f;
    const float h0 = x0 * 0.25f;
    const uint64_t lanes = __builtin_amdgcn_read_exec();
    for (int i = 0; i < iters; ++i) {
#define TVS_STEP(j)                                                                    \
    asm volatile("v_sub_f32_e32 %0, %0, %2\n\t"                                        \
                 "v_mul_f32_e32 %1, %3, %0\n\t"                                        \
                 "v_fmac_f32_e32 %1, %4, %0\n\t"                                       \
                 "v_fma_f32 %0, -%1, %3, %0\n\t"                                       \
                 "v_fmac_f32_e32 %1, %4, %0\n\t"                                       \
                 "v_subrev_f32_e32 %1, %3, %1\n\t"                                     \
                 "v_cmpx_gt_f32 vcc, %1, %2\n\t"                                       \
                 "v_add_f32_e32 %0, %4, %0\n\t"                                        \
                 "v_mov_b32_e32 %0, %1\n\t"                                            \
                 "s_mov_b64 exec, %5"                                                  \
                 : "+v"(x##j), "+v"(y##j)                                              \
                 : "v"(h0), "s"(c0), "s"(c1), "s"(lanes)                               \
                 : "vcc");
#define TVS_STEP3(j)                                                                   \
    asm volatile("v_sub_f32_e32 %0, %0, %2\n\t"                                        \
                 "v_mul_f32_e32 %1, %3, %0\n\t"                                        \
                 "v_fmac_f32_e32 %1, %4, %0\n\t"                                       \
                 "v_subrev_f32_e32 %1, %3, %1\n\t"                                     \
                 "v_cmpx_gt_f32 vcc, %1, %2\n\t"                                       \
                 "v_add_f32_e32 %0, %4, %0\n\t"                                        \
                 "v_mov_b32_e32 %0, %1\n\t"                                            \
                 "s_mov_b64 exec, %5"                                                  \
                 : "+v"(x##j), "+v"(y##j)                                              \
                 : "v"(h0), "s"(c0), "s"(c1), "s"(lanes)                               \
                 : "vcc");
#define TVS_FMA(j) asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(x##j) : "v"(y##j), "s"(c0));
#define TVS_ADD(j) asm volatile("v_add_f32_e32 %0, %1, %0" : "+v"(x##j) : "s"(c1));
        if (MODE == 0) { TVS_REP8(TVS_STEP) }
        if (MODE == 1) { TVS_REP8(TVS_FMA) }
        if (MODE == 2) { TVS_REP8(TVS_ADD) }
        if (MODE == 3) { TVS_REP8(TVS_STEP3) }
    }
    out[blockIdx.x * kBlock + threadIdx.x] =
        x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7 + y0 + y1 + y2 + y3 + y4 + y5 + y6 + y7;
}

template <int MODE>
int run(float *out, int blocks, int iters, int simds, double valu_per_iter, const char *name) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    hipLaunchKernelGGL(issue_kernel<MODE>, dim3(blocks), dim3(kBlock), 0, 0, out, iters,
                       0.999f, 0.5f);
    if (hipEventRecord(e0) != hipSuccess) return 1;
    hipLaunchKernelGGL(issue_kernel<MODE>, dim3(blocks), dim3(kBlock), 0, 0, out, iters,
                       0.999f, 0.5f);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    const double waves = blocks * (kBlock / 64.0);
    const double per_simd = waves * iters * valu_per_iter / simds;
    std::printf("{\"probe\": \"%s\", \"blocks\": %d, \"ms\": %.4f, "
                "\"ns_per_wave_valu_per_simd\": %.4f}\n",
                name, blocks, ms, ms * 1e6 / per_simd);
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
    const int simds = 4 * cus;
    const int blocks = cus * 8;   // 32 waves per CU: eight per SIMD
    float *out = nullptr;
    if (hipMalloc(&out, static_cast<size_t>(blocks) * kBlock * sizeof(float)) != hipSuccess)
        return 1;
    int rc = run<0>(out, blocks, 20000, simds, 8 * 9, "walk step (9 VALU + 1 SALU)");
    rc |= run<3>(out, blocks, 20000, simds, 8 * 7, "walk step, short divide (7 VALU + 1 SALU)");
    rc |= run<1>(out, blocks, 20000, simds, 8, "independent v_fma_f32");
    rc |= run<2>(out, blocks, 20000, simds, 8, "independent v_add_f32");
    (void)hipFree(out);
    return rc;
}
