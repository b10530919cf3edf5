// Ramp filter as a Toeplitz GEMM on MFMA (Mgfbp.ConvolveSgmAndKernel,
// run_mgfbp.py:1033-1046):  OUT[r, k] = scale * sum_m IN[r, m] T[m, k],
// T[m, k] = taps[k + n-1 - m].
//
// Workgroup tile: 128 rows x 128 output columns, four wavefronts of 64 x 64,
// each a 4 x 4 grid of v_mfma_f32_16x16x32_bf16 accumulators (FP32).  Per K
// step of 32:
//   * A (128 rows x 32 samples, FP32) is read from HBM, rounded to BF16 (hi)
//     and, for the split method, the remainder x - hi rounded again (lo), and
//     stored in LDS ([row][k], rows padded to 80 bytes);
//   * B is never materialised: every B fragment of the tile is 8 consecutive
//     entries of the reversed tap window R[i] = taps[top - i] (159 taps for a
//     128 x 32 tile), kept in LDS as 8 copies shifted by 0..7 so each
//     fragment is one aligned 16-byte LDS read.
// BF16: hi x hi.  BF16X3: hi x hi + hi x lo + lo x hi (relative product error
// about 2^-16 instead of 2^-8).
//
// MFMA operand layout (16x16x32, 64 lanes): lane l holds A[l % 16][8(l/16) .. +7],
// B[8(l/16) .. +7][l % 16]; the 4 accumulators of lane l are
// D[4(l/16) .. +3][l % 16].
#include "mct_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 128, TN = 128, TK = 32;
constexpr int APAD = 40;   // LDS row stride of the A tile, in BF16 elements (80 B)
constexpr int WLEN = 160;  // entries of one shifted window copy (needs 159 + alignment)

__device__ __forceinline__ __bf16 to_bf16(float f) {
    return __builtin_bit_cast(__bf16, mct::f32_to_bf16_rn(f));
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void ramp_mfma_kernel(const float *__restrict__ in,
                                                        float *__restrict__ out, int64_t rows, int n,
                                                        int npad, const __bf16 *__restrict__ taps_hi,
                                                        const __bf16 *__restrict__ taps_lo,
                                                        float scale) {
    __shared__ __attribute__((aligned(16))) __bf16 a_hi[TM * APAD];
    __shared__ __attribute__((aligned(16))) __bf16 a_lo[SPLIT ? TM * APAD : 8];
    __shared__ __attribute__((aligned(16))) __bf16 w_hi[8 * WLEN];
    __shared__ __attribute__((aligned(16))) __bf16 w_lo[SPLIT ? 8 * WLEN : 8];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t r0 = (int64_t)blockIdx.y * TM;
    const int k0 = blockIdx.x * TN;
    const int ntaps = 2 * n - 1;

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // A staging: thread t covers row t/2, samples 16*(t%2) .. +15 of the K step
    const int arow = tid >> 1, acol = (tid & 1) * 16;
    const int64_t grow = r0 + arow;

    for (int m0 = 0; m0 < npad; m0 += TK) {
        {
            float x[16];
            const bool rok = grow < rows;
            const float *src = in + grow * n + m0 + acol;
            if (rok && m0 + acol + 16 <= n && ((((uintptr_t)src) & 15) == 0)) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float4 v = reinterpret_cast<const float4 *>(src)[q];
                    x[4 * q] = v.x;
                    x[4 * q + 1] = v.y;
                    x[4 * q + 2] = v.z;
                    x[4 * q + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 16; ++q) x[q] = (rok && m0 + acol + q < n) ? src[q] : 0.f;
            }
            bf16x8 h0, h1, l0, l1;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                __bf16 a = to_bf16(x[q]), b = to_bf16(x[q + 8]);
                h0[q] = a;
                h1[q] = b;
                if (SPLIT) {
                    l0[q] = to_bf16(x[q] - (float)a);
                    l1[q] = to_bf16(x[q + 8] - (float)b);
                }
            }
            *reinterpret_cast<bf16x8 *>(&a_hi[arow * APAD + acol]) = h0;
            *reinterpret_cast<bf16x8 *>(&a_hi[arow * APAD + acol + 8]) = h1;
            if (SPLIT) {
                *reinterpret_cast<bf16x8 *>(&a_lo[arow * APAD + acol]) = l0;
                *reinterpret_cast<bf16x8 *>(&a_lo[arow * APAD + acol + 8]) = l1;
            }
        }
        // reversed tap window, 8 shifted copies: w[c][t] = taps[top - (t + c)]
        const int top = k0 - m0 + (TN - 1) + n - 1;
        for (int e = tid; e < 8 * WLEN; e += 256) {
            const int c = e / WLEN, t = e - c * WLEN;
            const int idx = top - (t + c);
            const bool ok = idx >= 0 && idx < ntaps;
            w_hi[e] = ok ? taps_hi[idx] : __builtin_bit_cast(__bf16, (uint16_t)0);
            if (SPLIT) w_lo[e] = ok ? taps_lo[idx] : __builtin_bit_cast(__bf16, (uint16_t)0);
        }
        __syncthreads();
        bf16x8 ah[4], al[4], bh[4], bl[4];
        const int kb = (lane >> 4) * 8;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const int row = wr * 64 + rb * 16 + (lane & 15);
            ah[rb] = *reinterpret_cast<const bf16x8 *>(&a_hi[row * APAD + kb]);
            if (SPLIT) al[rb] = *reinterpret_cast<const bf16x8 *>(&a_lo[row * APAD + kb]);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int jj = wc * 64 + cb * 16 + (lane & 15);
            const int s = (TN - 1) - jj + kb;  // B[kb + e][jj] = R[s + e]
            const int c = s & 7;
            bh[cb] = *reinterpret_cast<const bf16x8 *>(&w_hi[c * WLEN + (s - c)]);
            if (SPLIT) bl[cb] = *reinterpret_cast<const bf16x8 *>(&w_lo[c * WLEN + (s - c)]);
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bh[cb], acc[rb][cb], 0, 0, 0);
                if (SPLIT) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bl[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], bh[cb], acc[rb][cb], 0, 0, 0);
                }
            }
        __syncthreads();
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int col = k0 + wc * 64 + cb * 16 + (lane & 15);
            if (col >= n) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t r = r0 + wr * 64 + rb * 16 + (lane >> 4) * 4 + i;
                if (r < rows) out[r * n + col] = acc[rb][cb][i] * scale;
            }
        }
}

uint16_t bf16_bits_rn(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf16_value(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

using mct::fnv1a;

}  // namespace

namespace mct {

int filter_rows_mfma(mct_ctx *ctx, const float *in, float *out, int64_t n_rows, int n,
                     const float *taps, float scale, int method) {
    MCT_CHECK(method == MCT_FILTER_BF16_MFMA || method == MCT_FILTER_BF16X3_MFMA,
              "unknown filter method %d", method);
    const int ntaps = 2 * n - 1;
    // split taps into BF16 hi + lo once per tap set (cached in the context)
    const uint64_t h = fnv1a(taps, sizeof(float) * (size_t)ntaps, (uint64_t)n);
    if (!ctx->toep || ctx->toep_n != n || ctx->toep_hash != h) {
        std::vector<uint16_t> hl(2 * (size_t)ntaps);
        for (int i = 0; i < ntaps; ++i) {
            uint16_t hi = bf16_bits_rn(taps[i]);
            hl[i] = hi;
            hl[ntaps + i] = bf16_bits_rn(taps[i] - bf16_value(hi));
        }
        const int64_t bytes = (int64_t)hl.size() * 2;
        if (ctx->toep_bytes < bytes) {
            if (ctx->toep) {
                MCT_HIP(hipStreamSynchronize(ctx->stream));
                MCT_HIP(hipFree(ctx->toep));
                ctx->toep = nullptr;
                ctx->toep_bytes = 0;
            }
            MCT_HIP(hipMalloc(&ctx->toep, bytes));
            ctx->toep_bytes = bytes;
        }
        MCT_HIP(hipMemcpyAsync(ctx->toep, hl.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
        MCT_HIP(hipStreamSynchronize(ctx->stream));  // hl is released on return
        ctx->toep_n = n;
        ctx->toep_hash = h;
    }
    const __bf16 *thi = reinterpret_cast<const __bf16 *>(ctx->toep);
    const __bf16 *tlo = thi + ntaps;
    const int npad = (n + TK - 1) / TK * TK;
    const int64_t row_tiles = (n_rows + TM - 1) / TM;
    MCT_CHECK(row_tiles <= 65535, "too many rows for one launch (%lld)", (long long)n_rows);
    dim3 grid((n + TN - 1) / TN, (unsigned)row_tiles);
    if (method == MCT_FILTER_BF16X3_MFMA)
        hipLaunchKernelGGL(ramp_mfma_kernel<true>, grid, dim3(256), 0, ctx->stream, in, out, n_rows, n,
                           npad, thi, tlo, scale);
    else
        hipLaunchKernelGGL(ramp_mfma_kernel<false>, grid, dim3(256), 0, ctx->stream, in, out, n_rows,
                           n, npad, thi, tlo, scale);
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // namespace mct
