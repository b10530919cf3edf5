// VALU issue-rate probe for gfx950 (a measurement tool, not part of the
// library): each kernel runs one instruction form, or a fixed mix, on 16
// independent registers per thread (ALU only, no memory access besides one
// store per thread), over 2048 workgroups of 256 threads.  The wave64
// instruction count per launch is known exactly, so
//     rate = wave instructions / (CUs x launch time)
// is the issue rate per CU; with the launch's busy cycles from rocprofv3
// (GRBM_GUI_ACTIVE) it becomes instructions per CU per clock.  Two VALU
// instructions may issue in one SIMD cycle (SQ_ACTIVE_INST_VALU2 counts
// those cycles): the probe shows which forms can, and what the cone
// back-projector's slice-loop mix reaches without its LDS reads.
//
// Usage: issue_rate_probe  (prints one JSON line per kernel)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                \
        }                                                                            \
    } while (0)

namespace {

constexpr int kIters = 1024;
constexpr int kRegs = 16;
constexpr int kBlocks = 2048;

// __VA_ARGS__: the instructions of one register i; INSTS: their count.
#define PROBE(name, INSTS, ...)                                                               \
    constexpr int name##_insts = INSTS;                                                     \
    __global__ __launch_bounds__(256) void name(float *out, float s, uint32_t si) {         \
        float a[kRegs], b[kRegs];                                                           \
        for (int i = 0; i < kRegs; ++i) {                                                   \
            a[i] = threadIdx.x * 0.001f + i;                                                \
            b[i] = s * (0.5f + i * 0.01f) + threadIdx.x * 1e-9f;                            \
        }                                                                                   \
        for (int it = 0; it < kIters; ++it) {                                               \
            _Pragma("unroll") for (int i = 0; i < kRegs; ++i) { __VA_ARGS__; }              \
        }                                                                                   \
        float r = 0.f;                                                                      \
        for (int i = 0; i < kRegs; ++i) r += a[i] + b[i];                                   \
        out[blockIdx.x * 256 + threadIdx.x] = r + s + (float)si;                            \
    }

#define V(i) b[(i) % kRegs]
PROBE(fma_vgpr, 1, asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(a[i]) : "v"(V(i)), "v"(V(i + 1))))
PROBE(fma_literal, 1, asm volatile("v_fmamk_f32 %0, %1, 0x3faf5c29, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(fma_sgpr, 1, asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(a[i]) : "s"(s), "v"(V(i))))
PROBE(add_f32, 1, asm volatile("v_add_f32 %0, %1, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(mul_f32, 1, asm volatile("v_mul_f32 %0, %1, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(add_u32, 1, asm volatile("v_add_u32 %0, %1, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(and_b32, 1, asm volatile("v_and_b32 %0, %1, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(min_u32, 1, asm volatile("v_min_u32 %0, %1, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(lshl_add_u32, 1, asm volatile("v_lshl_add_u32 %0, %1, 4, %0" : "+v"(a[i]) : "v"(V(i))))
PROBE(mad_u32_u24, 1, asm volatile("v_mad_u32_u24 %0, %1, %2, %0" : "+v"(a[i]) : "v"(V(i)), "v"(V(i + 1))))
PROBE(cvt_u32_f32, 1, asm volatile("v_cvt_u32_f32 %0, %0" : "+v"(a[i])))
PROBE(fract_f32, 1, asm volatile("v_fract_f32 %0, %0" : "+v"(a[i])))
// one fract (single issue) with one or three VGPR-only FMAs
PROBE(fract_1fma, 2,
      asm volatile("v_fract_f32 %0, %0" : "+v"(a[i]));
      asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(b[i]) : "v"(a[(i + 8) % kRegs]), "v"(V(i + 1))))
PROBE(fract_3fma, 4,
      asm volatile("v_fract_f32 %0, %0" : "+v"(a[i]));
      asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(b[i]) : "v"(a[(i + 8) % kRegs]), "v"(V(i + 1)));
      asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(b[(i + 4) % kRegs]) : "v"(a[(i + 9) % kRegs]), "v"(V(i + 5)));
      asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(b[(i + 12) % kRegs]) : "v"(a[(i + 10) % kRegs]), "v"(V(i + 13))))
// the cone back-projector's slice-loop instructions per update, without the
// LDS read, written as the kernel writes them (the compiler emits the same
// forms: a literal FMA for the row coordinate, convert, fract, an address
// multiply-add with an SGPR stride, three bilinear FMAs and the
// accumulation FMA)
PROBE(bp_slice_mix, 8,
      const float tk = fmaf(3.0f, V(i), a[i]);
      const uint32_t vf = (uint32_t)tk;
      const float rv = __builtin_amdgcn_fractf(tk);
      const float q = __uint_as_float(__umul24(vf, si) + __float_as_uint(V(i + 2)));
      const float x = fmaf(rv, V(i + 3), q);
      const float y = fmaf(rv, V(i + 4), V(i + 5));
      a[i] = fmaf(V(i + 6), fmaf(rv, y, x), a[i]))
// dependent FMAs: one chain per wave (each instruction uses the result of
// the one before it), and two or four chains interleaved
PROBE(fma_chain1, 1, a[0] = fmaf(a[0], V(i), V(i + 1)))
PROBE(fma_chain2, 1, a[i & 1] = fmaf(a[i & 1], V(i), V(i + 1)))
PROBE(fma_chain4, 1, a[i & 3] = fmaf(a[i & 3], V(i), V(i + 1)))
#undef V

struct Probe {
    void (*fn)(float *, float, uint32_t);
    const char *name;
    int insts;
};
#define P(n) {n, #n, n##_insts}
const Probe kProbes[] = {P(fma_vgpr),   P(fma_literal), P(fma_sgpr),     P(add_f32),     P(mul_f32),
                         P(add_u32),    P(and_b32),     P(min_u32),      P(lshl_add_u32), P(mad_u32_u24),
                         P(cvt_u32_f32), P(fract_f32),  P(fract_1fma),   P(fract_3fma),  P(bp_slice_mix),
                         P(fma_chain1),  P(fma_chain2),  P(fma_chain4)};
#undef P

}  // namespace

int main() {
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    float *out;
    CHECK(hipMalloc(&out, (size_t)kBlocks * 256 * sizeof(float)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int rep = 0; rep < 2; ++rep) {  // the first pass warms up
        for (const Probe &p : kProbes) {
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(p.fn, dim3(kBlocks), dim3(256), 0, 0, out, 0.999f, 5u);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            const double waves = (double)kBlocks * 4;
            const double insts = waves * kIters * kRegs * p.insts;  // the timed loop's wave instructions
            if (rep)
                std::printf("{\"probe\": \"%s\", \"ms\": %.4f, \"wave_insts\": %.0f, "
                            "\"wave_insts_per_cu_per_ns\": %.3f}\n",
                            p.name, ms, insts, insts / cus / (ms * 1e6));
        }
    }
    CHECK(hipFree(out));
    return 0;
}
