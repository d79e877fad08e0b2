// pair_probe -- which vector instructions of the walk's step issue together.
//
// A SIMD of this part can issue two vector instructions in one quad-cycle
// (the SQ_ACTIVE_INST_VALU2 counter counts those quad-cycles).  Each probe
// kernel runs, at eight waves per SIMD on every SIMD, 20 000 iterations of
// eight independent chains of one instruction pair (x = A(x), y = B(x, y))
// out of registers, with no memory access, and reports cycles (at 2.4 GHz)
// per wave-instruction per SIMD: about 4.35 when the pair never issues
// together, down to about 2.5 when it always does.  The pairs are the
// walk's own instructions: adds and multiplies with a scalar operand (the
// per-step constants), the same with vector operands only, max, compare,
// the exec-masking compare, and the SDWA conversion of a 16-bit half.
//
// Build: make -C total-viewsheds_amd probe.  Run: total-viewsheds_amd/pair_probe
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlock = 256;
constexpr int kIters = 20000;

// x: first operand chain, y: second; s0, s1: scalar constants; e: the exec
// mask to restore after a v_cmpx.
#define TVS_PAIR_KERNEL(NAME, TEXT)                                                          \
    __global__ __launch_bounds__(kBlock) void NAME(float *out, float s0, float s1) {         \
        float x[8], y[8];                                                                      \
        const uint64_t e = __builtin_amdgcn_read_exec();                                       \
        for (int j = 0; j < 8; ++j) {                                                          \
            x[j] = threadIdx.x * 1e-3f + j;                                                    \
            y[j] = x[j] * 0.5f;                                                                \
        }                                                                                      \
        for (int i = 0; i < kIters; ++i) {                                                     \
            _Pragma("unroll") for (int j = 0; j < 8; ++j)                                      \
                asm volatile(TEXT : "+v"(x[j]), "+v"(y[j]) : "s"(s0), "s"(s1), "s"(e) : "vcc");  \
        }                                                                                      \
        float r = 0.0f;                                                                        \
        for (int j = 0; j < 8; ++j) r += x[j] + y[j];                                          \
        out[blockIdx.x * kBlock + threadIdx.x] = r;                                            \
    }

TVS_PAIR_KERNEL(add_s__add_s, "v_add_f32_e32 %0, %3, %0\n\tv_add_f32_e32 %1, %3, %1")
TVS_PAIR_KERNEL(mul_s__fmac_s, "v_mul_f32_e32 %0, %2, %0\n\tv_fmac_f32_e32 %1, %2, %0")
TVS_PAIR_KERNEL(add_v__add_s, "v_add_f32_e32 %0, %1, %0\n\tv_add_f32_e32 %1, %3, %1")
TVS_PAIR_KERNEL(mul_v__mul_s, "v_mul_f32_e32 %0, %1, %0\n\tv_mul_f32_e32 %1, %2, %1")
TVS_PAIR_KERNEL(add_s__max, "v_add_f32_e32 %0, %3, %0\n\tv_max_f32_e32 %1, %0, %1")
TVS_PAIR_KERNEL(fmac_s__max, "v_fmac_f32_e32 %0, %2, %1\n\tv_max_f32_e32 %1, %0, %1")
TVS_PAIR_KERNEL(max__max, "v_max_f32_e32 %0, %1, %0\n\tv_max_f32_e32 %1, %0, %1")
TVS_PAIR_KERNEL(cmp__add_s, "v_cmp_gt_f32_e32 vcc, %0, %1\n\tv_add_f32_e32 %1, %3, %1")
TVS_PAIR_KERNEL(cmp__max, "v_cmp_gt_f32_e32 vcc, %0, %1\n\tv_max_f32_e32 %1, %0, %1")
TVS_PAIR_KERNEL(cmpx__add_s,
                "v_cmpx_gt_f32_e32 vcc, %0, %1\n\ts_mov_b64 exec, %4\n\tv_add_f32_e32 %1, %3, %1")
TVS_PAIR_KERNEL(cvt_sdwa__add_s,
                "v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:WORD_1\n\tv_add_f32_e32 %1, %3, %1")
TVS_PAIR_KERNEL(cvt_sdwa__max,
                "v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:WORD_1\n\tv_max_f32_e32 %1, %0, %1")
TVS_PAIR_KERNEL(cvt__add_s, "v_cvt_f32_i32_e32 %0, %1\n\tv_add_f32_e32 %1, %3, %1")
TVS_PAIR_KERNEL(sub_v__max, "v_sub_f32_e32 %0, %0, %1\n\tv_max_f32_e32 %1, %0, %1")
TVS_PAIR_KERNEL(mov__max, "v_mov_b32_e32 %0, %1\n\tv_max_f32_e32 %1, %0, %1")

template <typename K>
int run(K kernel, float *out, int blocks, int simds, const char *name) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, 0, out, 0.999f, 0.5f);
    if (hipEventRecord(e0) != hipSuccess) return 1;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, 0, out, 0.999f, 0.5f);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    // Two vector instructions per chain and iteration, eight chains.
    const double per_simd = blocks * (kBlock / 64.0) * kIters * 16.0 / simds;
    std::printf("{\"pair\": \"%s\", \"ms\": %.4f, \"cycles_per_wave_valu_per_simd\": %.3f}\n",
                name, ms, ms * 1e6 / per_simd * 2.4);
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
    int rc = 0;
    rc |= run(add_s__add_s, out, blocks, simds, "add(s) + add(s)");
    rc |= run(mul_s__fmac_s, out, blocks, simds, "mul(s) + fmac(s)");
    rc |= run(add_v__add_s, out, blocks, simds, "add(v) + add(s)");
    rc |= run(mul_v__mul_s, out, blocks, simds, "mul(v) + mul(s)");
    rc |= run(add_s__max, out, blocks, simds, "add(s) + max");
    rc |= run(fmac_s__max, out, blocks, simds, "fmac(s) + max");
    rc |= run(max__max, out, blocks, simds, "max + max");
    rc |= run(cmp__add_s, out, blocks, simds, "cmp + add(s)");
    rc |= run(cmp__max, out, blocks, simds, "cmp + max");
    rc |= run(cmpx__add_s, out, blocks, simds, "cmpx, exec restore + add(s)");
    rc |= run(cvt_sdwa__add_s, out, blocks, simds, "cvt sdwa + add(s)");
    rc |= run(cvt_sdwa__max, out, blocks, simds, "cvt sdwa + max");
    rc |= run(cvt__add_s, out, blocks, simds, "cvt + add(s)");
    rc |= run(sub_v__max, out, blocks, simds, "sub(v) + max");
    rc |= run(mov__max, out, blocks, simds, "mov + max");
    (void)hipFree(out);
    return rc;
}
