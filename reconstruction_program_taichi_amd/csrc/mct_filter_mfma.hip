// Ramp filter as a Toeplitz GEMM on MFMA (Mgfbp.ConvolveSgmAndKernel,
// run_mgfbp.py:1033-1046):  OUT[r, k] = scale * sum_m IN[r, m] taps[k + n-1 - m].
//
// Difference form.  The ramp taps sum to almost zero, so for a smooth row the
// products x_m h_{k-m} cancel to an output far smaller than themselves, and
// any rounding of the operands is amplified by that ratio (weighted
// projections of ~2e3 filter to ~1).  Summation by parts removes it: with
// g_j = h_{-(n-1)} + ... + h_j the running sum of the taps (g_{-n} = 0) and
// D_m = x_m - x_{m-1} the first difference of the row (x_{-1} = x_n = 0),
//     sum_{m<n} x_m h_{k-m} = sum_{m<=n} D_m g_{k-m}
// exactly.  The MFMA operands are then D (small where the row is smooth) and
// g (bounded by about h_0 / 2), so the BF16 rounding scales with |D| instead
// of |x|.  The product is still Toeplitz: OUT[r, k] = scale * sum_{m<=n}
// D[r, m] G[k + n - m] with G[i] = taps[0] + ... + taps[i-1] (2n entries,
// G[0] = 0), over K = n + 1 samples.
//
// Two launches:
//   * split_bf16_kernel turns every input row once into BF16 planes of its
//     first differences, [rows][npad] with npad >= n + 1: hi = bf16(D) and,
//     for the split method, lo = bf16(D - hi), zero past n; with a WeighSpec
//     it first weighs each sample exactly as the weighting kernel does (same
//     FP32 operations), so weighting and filtering read the sinogram once;
//   * ramp_gemm_kernel: workgroup tile 128 rows x 128 output columns, four
//     wavefronts of 64 x 64, each a 4 x 4 grid of v_mfma_f32_16x16x32_bf16
//     accumulators (FP32).  Per K step of 32 the A tile (128 rows x 32
//     samples of each plane) is copied from the planes into one of two LDS
//     buffers; the copy for step s+1 is in flight while step s's MFMAs run.
//     B is never materialised: every B fragment is 8 consecutive entries of
//     the reversed window W[i] = G[T - i], T = k0 + 127 + n, kept in LDS as 8
//     copies shifted by 0..7 so each fragment is one aligned 16-byte LDS
//     read; a window covers KC samples of K and is restaged every KC/32
//     steps.  Workgroup ids are remapped so the 8 column tiles of a row tile
//     run on one XCD and share its L2.
// BF16: hi x hi.  BF16X3: hi x hi + hi x lo + lo x hi (relative product error
// about 2^-16 instead of 2^-8), in that order for every accumulator.
//
// MFMA operand layout (16x16x32, 64 lanes): lane l holds A[l % 16][8(l/16) .. +7],
// B[8(l/16) .. +7][l % 16]; the 4 accumulators of lane l are
// D[4(l/16) .. +3][l % 16].
#include "mct_internal.h"
#include "mct_weigh.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 128, TN = 128, TK = 32;
constexpr int APAD = 40;      // LDS row stride of an A tile, in BF16 elements (80 B)
constexpr int KC = 256;  // K samples covered by one staged tap window
// Entries per shifted copy: fragments start below KC - TK + TN + 24, plus 16
// entries of padding: a copy then spans 8 banks mod 64 more than a multiple
// of 64, so the 16 lanes of a fragment read (two groups of 8 copies, the
// groups 4 banks apart) land on 64 distinct banks.
constexpr int WLC = KC + TN + 16;

__device__ __forceinline__ __bf16 to_bf16(float f) {
    return __builtin_bit_cast(__bf16, mct::f32_to_bf16_rn(f));
}

// One thread per 8 samples of a row: the differences D[m0 .. m0+7] of
// in[r][m0-1 .. m0+7] (samples outside [0, n) are 0) -> hi/lo planes
// (WEIGH: row r is (detector row r / w.nviews, view r % w.nviews of the block)).
template <bool SPLIT, bool WEIGH>
__global__ __launch_bounds__(256) void split_bf16_kernel(const float *__restrict__ in, int64_t rows,
                                                         int n, int npad, __bf16 *__restrict__ hi,
                                                         __bf16 *__restrict__ lo, mct::WeighSpec w) {
    const int groups = npad / 8;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= rows * groups) return;
    const int64_t r = t / groups;
    const int m0 = (int)(t - r * groups) * 8;
    const float *src = in + r * n + m0;
    float x[8];
    float xp = (m0 >= 1 && m0 - 1 < n) ? src[-1] : 0.f;  // sample m0 - 1
    if (m0 + 8 <= n && ((((uintptr_t)src) & 15) == 0)) {
        const float4 a = reinterpret_cast<const float4 *>(src)[0];
        const float4 b = reinterpret_cast<const float4 *>(src)[1];
        x[0] = a.x, x[1] = a.y, x[2] = a.z, x[3] = a.w;
        x[4] = b.x, x[5] = b.y, x[6] = b.z, x[7] = b.w;
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = m0 + q < n ? src[q] : 0.f;
    }
    if (WEIGH) {
        const int s = (int)(r / w.nviews), i = (int)(r - (int64_t)s * w.nviews);
        float u[8];
        if (m0 + 8 <= n) {  // the positions array is 256-byte aligned and m0 a multiple of 8
            const float4 a = reinterpret_cast<const float4 *>(w.u + m0)[0];
            const float4 b = reinterpret_cast<const float4 *>(w.u + m0)[1];
            u[0] = a.x, u[1] = a.y, u[2] = a.z, u[3] = a.w;
            u[4] = b.x, u[5] = b.y, u[6] = b.z, u[7] = b.w;
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) u[q] = m0 + q < n ? w.u[m0 + q] : 0.f;
        }
        const float v = w.v[s];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (m0 + q < n)
                x[q] = mct::weigh(x[q], u[q], v, w.view0 + i, w.ang, w.sdd, w.scan, w.short_scan,
                                  w.curved, w.bh, w.bh_order);
        if (m0 >= 1 && m0 - 1 < n)
            xp = mct::weigh(xp, w.u[m0 - 1], v, w.view0 + i, w.ang, w.sdd, w.scan, w.short_scan,
                            w.curved, w.bh, w.bh_order);
    }
    bf16x8 h, l;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float d = x[q] - (q ? x[q - 1] : xp);
        h[q] = to_bf16(d);
        if (SPLIT) l[q] = to_bf16(d - (float)h[q]);
    }
    *reinterpret_cast<bf16x8 *>(hi + r * npad + m0) = h;
    if (SPLIT) *reinterpret_cast<bf16x8 *>(lo + r * npad + m0) = l;
}

// Lane l's value from lane (l & ~15) | ((l + 1) & 15): the next column of a
// 16 x 16 accumulator block (DPP row rotate by 15 within each row of 16).
__device__ __forceinline__ float next_col(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x12F, 0xF, 0xF, false));
}

// ROWS_OUT: instead of the FP32 rows, the epilogue writes the back-projector's
// ROWS taps of every filtered element, {y[c], y[c+1] - y[c]} (FP32), to
// out[view][row + 1][c] (row = r / views, view = r % views of GEMM row r;
// out has det_rows + 2 rows of n - 1 entries per view).  A column's right
// neighbour is the next lane of its accumulator block, or the first of the
// next block; the last column of each wavefront's 64 has its neighbour in
// another wavefront, and rows_fixup_kernel completes its difference.
template <bool SPLIT, bool ROWS_OUT>
__global__ __launch_bounds__(256) void ramp_gemm_kernel(const __bf16 *__restrict__ plane_hi,
                                                        const __bf16 *__restrict__ plane_lo,
                                                        float *__restrict__ out, int64_t rows, int n,
                                                        int npad, int col_tiles, int nxcd,
                                                        const __bf16 *__restrict__ taps_hi,
                                                        const __bf16 *__restrict__ taps_lo, float scale,
                                                        int views, int det_rows) {
    __shared__ __attribute__((aligned(16))) __bf16 a_hi[2][TM * APAD];
    __shared__ __attribute__((aligned(16))) __bf16 a_lo[2][SPLIT ? TM * APAD : 8];
    __shared__ __attribute__((aligned(16))) __bf16 w_hi[8 * WLC];
    __shared__ __attribute__((aligned(16))) __bf16 w_lo[SPLIT ? 8 * WLC : 8];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int tile = mct::xcd_remap((int)blockIdx.x, (int)gridDim.x, nxcd);
    const int64_t r0 = (int64_t)(tile / col_tiles) * TM;
    const int k0 = (tile % col_tiles) * TN;
    const int ntaps = 2 * n;  // G[0 .. 2n-1]
    const int T = k0 + (TN - 1) + n;

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // A copy: thread t moves row t/2, samples 16 (t % 2) .. +15 of a K step
    // (two 16-byte pieces of each plane)
    const int arow = tid >> 1, acol = (tid & 1) * 16;
    const int64_t grow = r0 + arow;
    const bool rok = grow < rows;
    const __bf16 *ph = plane_hi + (rok ? grow : 0) * npad + acol;
    const __bf16 *pl = plane_lo + (rok ? grow : 0) * npad + acol;
    const bf16x8 zero8 = {};
    bf16x8 sh0 = zero8, sh1 = zero8, sl0 = zero8, sl1 = zero8;
    auto load_a = [&](int m0) {
        if (rok) {
            sh0 = *reinterpret_cast<const bf16x8 *>(ph + m0);
            sh1 = *reinterpret_cast<const bf16x8 *>(ph + m0 + 8);
            if (SPLIT) {
                sl0 = *reinterpret_cast<const bf16x8 *>(pl + m0);
                sl1 = *reinterpret_cast<const bf16x8 *>(pl + m0 + 8);
            }
        }
    };
    auto store_a = [&](int b) {
        *reinterpret_cast<bf16x8 *>(&a_hi[b][arow * APAD + acol]) = sh0;
        *reinterpret_cast<bf16x8 *>(&a_hi[b][arow * APAD + acol + 8]) = sh1;
        if (SPLIT) {
            *reinterpret_cast<bf16x8 *>(&a_lo[b][arow * APAD + acol]) = sl0;
            *reinterpret_cast<bf16x8 *>(&a_lo[b][arow * APAD + acol + 8]) = sl1;
        }
    };
    // w[c][t] = W[mb + t + c]
    auto load_window = [&](int mb) {
        for (int e = tid; e < 8 * WLC; e += 256) {
            const int c = e / WLC, t = e - c * WLC;
            const int idx = T - mb - t - c;
            const bool ok = idx >= 0 && idx < ntaps;
            w_hi[e] = ok ? taps_hi[idx] : __builtin_bit_cast(__bf16, (uint16_t)0);
            if (SPLIT) w_lo[e] = ok ? taps_lo[idx] : __builtin_bit_cast(__bf16, (uint16_t)0);
        }
    };

    load_a(0);
    store_a(0);
    const int nsteps = npad / TK;
    const int kb = (lane >> 4) * 8;
    for (int step = 0; step < nsteps; ++step) {
        const int m0 = step * TK, cur = step & 1;
        const int mw = m0 % KC;  // position in the staged window
        if (mw == 0) {
            if (step) __syncthreads();  // every wave is done with the previous window
            load_window(m0);
        }
        __syncthreads();  // A buffer `cur` and the window are in LDS
        if (step + 1 < nsteps) load_a(m0 + TK);
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const int row = wr * 64 + rb * 16 + (lane & 15);
            ah[rb] = *reinterpret_cast<const bf16x8 *>(&a_hi[cur][row * APAD + kb]);
            if (SPLIT) al[rb] = *reinterpret_cast<const bf16x8 *>(&a_lo[cur][row * APAD + kb]);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int jj = wc * 64 + cb * 16 + (lane & 15);
            const int p = mw + (TN - 1) - jj + kb;  // B[kb + e][jj] = W[m0 + p' + e], p' = p - mw
            const int c = p & 7;
            bh[cb] = *reinterpret_cast<const bf16x8 *>(&w_hi[c * WLC + (p - c)]);
            if (SPLIT) bl[cb] = *reinterpret_cast<const bf16x8 *>(&w_lo[c * WLC + (p - c)]);
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
        if (step + 1 < nsteps) store_a(cur ^ 1);  // no wave reads buffer cur^1 after this step's barrier
    }
    if constexpr (ROWS_OUT) {
        float2 *po = reinterpret_cast<float2 *>(out);
        const int nc = n - 1;
        const bool last_lane = (lane & 15) == 15;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t r = r0 + wr * 64 + rb * 16 + (lane >> 4) * 4 + i;
                const int64_t dr = r / views, vw = r - dr * views;
                const int64_t base = (vw * (det_rows + 2) + dr + 1) * nc;
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    const float y = acc[rb][cb][i] * scale;
                    const float ys = next_col(y);
                    const float yn = cb < 3 ? next_col(acc[rb][cb < 3 ? cb + 1 : cb][i] * scale) : 0.f;
                    const float nb = last_lane ? yn : ys;
                    const int col = k0 + wc * 64 + cb * 16 + (lane & 15);
                    if (r < rows && col < nc) po[base + col] = make_float2(y, nb - y);
                }
            }
    } else {
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
}

// Completes the ROWS taps written by ramp_gemm_kernel<., true>: the
// difference of every column 64 t + 63 < n - 1 (its neighbour was in another
// wavefront; both values are final now), and the zero rows 0 and det_rows + 1
// of every view.
__global__ __launch_bounds__(256) void rows_fixup_kernel(float2 *__restrict__ po, int64_t rows, int n,
                                                         int views, int det_rows) {
    const int nc = n - 1;
    // elements c = 64 t + 63 <= nc - 1 (their neighbour c + 1 <= nc - 1 is an
    // element too: the caller keeps nc from being a multiple of 64)
    const int nb = nc >= 64 ? (nc - 64) / 64 + 1 : 0;
    const int64_t nfix = rows * nb, nzero = (int64_t)views * 2 * nc;
    const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id < nfix) {
        const int64_t r = id / nb;
        const int c = (int)(id - r * nb) * 64 + 63;
        const int64_t dr = r / views, vw = r - dr * views;
        float2 *e = po + (vw * (det_rows + 2) + dr + 1) * nc + c;
        e[0].y = e[1].x - e[0].x;
    } else if (id < nfix + nzero) {
        const int64_t k = id - nfix;
        const int64_t vw = k / (2 * nc);
        const int rest = (int)(k - vw * 2 * nc);
        const int s = rest < nc ? 0 : det_rows + 1, c = rest < nc ? rest : rest - nc;
        po[(vw * (det_rows + 2) + s) * nc + c] = make_float2(0.f, 0.f);
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

// w != NULL: the input rows are an unweighted [n_v][w->nviews][n] block,
// weighed in the split pass (n_rows = n_v * w->nviews).  rows_views > 0:
// out receives the FP32 ROWS taps of the [n_rows / rows_views][rows_views][n]
// rows instead (views major, rows padded by a zero row either side; n - 1
// must not be a multiple of 64).
int filter_rows_mfma(mct_ctx *ctx, const float *in, float *out, int64_t n_rows, int n,
                     const float *taps, float scale, int method, const WeighSpec *w, int rows_views) {
    MCT_CHECK(method == MCT_FILTER_BF16_MFMA || method == MCT_FILTER_BF16X3_MFMA,
              "unknown filter method %d", method);
    const int ntaps = 2 * n;  // running sums G[0 .. 2n-1] of the 2n-1 taps
    // running sums of the taps (FP64), split into BF16 hi + lo once per tap
    // set (cached in the context)
    const uint64_t h = fnv1a(taps, sizeof(float) * (size_t)(2 * n - 1), (uint64_t)n);
    if (!ctx->toep || ctx->toep_n != n || ctx->toep_hash != h) {
        if (ctx->frozen) return mct::refuse_frozen("the ramp tap panel (n)", ctx->toep_n, n);
        std::vector<uint16_t> hl(2 * (size_t)ntaps);
        double g = 0.0;
        for (int i = 0; i < ntaps; ++i) {
            if (i) g += (double)taps[i - 1];
            const float gf = (float)g;
            uint16_t hi = bf16_bits_rn(gf);
            hl[i] = hi;
            hl[ntaps + i] = bf16_bits_rn(gf - bf16_value(hi));
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
    const bool split = method == MCT_FILTER_BF16X3_MFMA;
    const int npad = (n + 1 + TK - 1) / TK * TK;  // K = n + 1 differences D_0 .. D_n
    const int64_t row_tiles = (n_rows + TM - 1) / TM;
    const int col_tiles = (n + TN - 1) / TN;
    MCT_CHECK(row_tiles * col_tiles <= INT32_MAX, "too many rows for one launch (%lld)",
              (long long)n_rows);
    // BF16 planes of the input rows (context scratch, kept between calls)
    const int64_t plane = n_rows * (int64_t)npad;
    const int64_t bytes = plane * 2 * (split ? 2 : 1);
    if (ctx->fsplit_bytes < bytes) {
        if (ctx->frozen) return mct::refuse_frozen("the filter input planes (bytes)", ctx->fsplit_bytes, bytes);
        if (ctx->fsplit) {
            MCT_HIP(hipStreamSynchronize(ctx->stream));
            MCT_HIP(hipFree(ctx->fsplit));
            ctx->fsplit = nullptr;
            ctx->fsplit_bytes = 0;
        }
        MCT_HIP(hipMalloc(&ctx->fsplit, bytes));
        ctx->fsplit_bytes = bytes;
    }
    __bf16 *phi = reinterpret_cast<__bf16 *>(ctx->fsplit);
    __bf16 *plo = split ? phi + plane : phi;
    const int64_t groups = n_rows * (npad / 8);
    const dim3 sgrid((unsigned)((groups + 255) / 256));
    const dim3 grid((unsigned)(row_tiles * col_tiles));
    const WeighSpec wv = w ? *w : WeighSpec{};
    const int rv = rows_views > 0 ? rows_views : 1;
    const int det_rows = (int)(n_rows / rv);
    if (rows_views > 0)
        MCT_CHECK(n >= 2 && (n - 1) % 64 != 0 && n_rows % rows_views == 0,
                  "ROWS output needs n - 1 not a multiple of 64 (n = %d)", n);
#define GEMM(SPLIT)                                                                              \
    do {                                                                                         \
        if (rows_views > 0)                                                                      \
            hipLaunchKernelGGL((ramp_gemm_kernel<SPLIT, true>), grid, dim3(256), 0, ctx->stream, \
                               phi, plo, out, n_rows, n, npad, col_tiles, ctx->xcd_count, thi, tlo, \
                               scale, rv, det_rows);                                             \
        else                                                                                     \
            hipLaunchKernelGGL((ramp_gemm_kernel<SPLIT, false>), grid, dim3(256), 0, ctx->stream, \
                               phi, plo, out, n_rows, n, npad, col_tiles, ctx->xcd_count, thi, tlo, \
                               scale, rv, det_rows);                                             \
    } while (0)
    if (split) {
        if (w)
            hipLaunchKernelGGL((split_bf16_kernel<true, true>), sgrid, dim3(256), 0, ctx->stream, in,
                               n_rows, n, npad, phi, plo, wv);
        else
            hipLaunchKernelGGL((split_bf16_kernel<true, false>), sgrid, dim3(256), 0, ctx->stream, in,
                               n_rows, n, npad, phi, plo, wv);
        GEMM(true);
    } else {
        if (w)
            hipLaunchKernelGGL((split_bf16_kernel<false, true>), sgrid, dim3(256), 0, ctx->stream, in,
                               n_rows, n, npad, phi, plo, wv);
        else
            hipLaunchKernelGGL((split_bf16_kernel<false, false>), sgrid, dim3(256), 0, ctx->stream, in,
                               n_rows, n, npad, phi, plo, wv);
        GEMM(false);
    }
#undef GEMM
    if (rows_views > 0) {
        const int nc = n - 1, nb = nc >= 64 ? (nc - 64) / 64 + 1 : 0;
        const int64_t work = n_rows * nb + (int64_t)rv * 2 * nc;
        hipLaunchKernelGGL(rows_fixup_kernel, dim3(mct::blocks_for(work, 256)), dim3(256), 0, ctx->stream,
                           reinterpret_cast<float2 *>(out), n_rows, n, rv, det_rows);
    }
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // namespace mct
