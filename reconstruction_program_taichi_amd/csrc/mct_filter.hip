// Ramp-filter row convolution (Mgfbp.ConvolveSgmAndKernel,
// run_mgfbp.py:1033-1046):
//   out[r, k] = scale * sum_{m=0}^{n-1} in[r, m] * taps[k + n-1 - m]
// i.e. each row times the n x n Toeplitz matrix T[m][k] = taps[k + n-1 - m].
//
// F32_DIRECT: FP32 FMAs in the reference's summation order (ascending m).  A
// workgroup owns RB = 4 rows, stored interleaved in LDS ([m][4]) so one
// broadcast 16-byte LDS read feeds the four rows.  A thread owns KB = 4
// consecutive outputs; their taps form a window that slides down by one tap
// per m, so each m costs one new tap read from LDS and 16 FMAs.
#include "mct_internal.h"
#include "mct_weigh.h"

namespace {

constexpr int RB = 4;  // rows per workgroup
constexpr int KB = 4;  // outputs per thread

__global__ __launch_bounds__(256) void ramp_direct_kernel(const float *__restrict__ in,
                                                          float *__restrict__ out, int64_t n_rows,
                                                          int n, const float *__restrict__ taps,
                                                          float scale) {
    extern __shared__ float smem[];
    float4 *rows4 = reinterpret_cast<float4 *>(smem);  // [n] x 4 rows
    float *tp = smem + 4 * (size_t)n;                    // [2n - 1] taps, then one pad
    const int64_t r0 = (int64_t)blockIdx.x * RB;
    const int nr = (int)((n_rows - r0) < RB ? (n_rows - r0) : RB);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        float4 v;
        v.x = in[(r0 + 0) * n + i];
        v.y = nr > 1 ? in[(r0 + 1) * n + i] : 0.f;
        v.z = nr > 2 ? in[(r0 + 2) * n + i] : 0.f;
        v.w = nr > 3 ? in[(r0 + 3) * n + i] : 0.f;
        rows4[i] = v;
    }
    for (int i = threadIdx.x; i < 2 * n - 1; i += blockDim.x) tp[i] = taps[i];
    if (threadIdx.x == 0) tp[2 * n - 1] = 0.f;
    __syncthreads();
    for (int k0 = threadIdx.x * KB; k0 < n; k0 += blockDim.x * KB) {
        float acc[RB][KB];
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int j = 0; j < KB; ++j) acc[r][j] = 0.f;
        // window w[j] = taps[k0 + j + n-1 - m]; indexes stay within [0, 2n-1]
        // (index 2n-1 is the zero pad, used only by outputs k >= n).
        float w[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            int t = k0 + j + n - 1;
            w[j] = t <= 2 * n - 1 ? tp[t] : 0.f;
        }
        for (int m = 0; m < n; ++m) {
            float4 x = rows4[m];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                acc[0][j] = fmaf(x.x, w[j], acc[0][j]);
                acc[1][j] = fmaf(x.y, w[j], acc[1][j]);
                acc[2][j] = fmaf(x.z, w[j], acc[2][j]);
                acc[3][j] = fmaf(x.w, w[j], acc[3][j]);
            }
#pragma unroll
            for (int j = KB - 1; j > 0; --j) w[j] = w[j - 1];
            int t = k0 + n - 2 - m;
            w[0] = t >= 0 ? tp[t] : 0.f;
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            if (r >= nr) break;
#pragma unroll
            for (int j = 0; j < KB; ++j)
                if (k0 + j < n) out[(r0 + r) * n + k0 + j] = acc[r][j] * scale;
        }
    }
}

}  // namespace

namespace mct {
int filter_rows_mfma(mct_ctx *ctx, const float *in, float *out, int64_t n_rows, int n,
                     const float *taps, float scale, int method, const WeighSpec *w, int rows_views);
}

namespace {

// Checks the arguments of a fused weighting + MFMA filter call and fills the
// weighting of views [view_begin, view_begin + view_count) of the plan.
int weigh_spec(mct_plan *plan, int32_t view_begin, int32_t view_count, const float *bh_coef,
               int32_t bh_order, int32_t method, mct::WeighSpec &w) {
    MCT_CHECK(method == MCT_FILTER_BF16_MFMA || method == MCT_FILTER_BF16X3_MFMA,
              "weighting is fused only into the MFMA filter methods (got %d)", method);
    MCT_CHECK(bh_order >= 0 && bh_order <= mct::kMaxBH, "bh_order must be in [0, %d]", mct::kMaxBH);
    MCT_CHECK(bh_order == 0 || bh_coef, "bh_coef is NULL");
    const mct_geometry &g = plan->g;
    MCT_CHECK(view_begin >= 0 && view_count >= 1 && view_count <= g.n_views - view_begin,
              "views [%d, %d + %d) outside the scan's %d views", view_begin, view_begin, view_count,
              g.n_views);
    MCT_CHECK(g.n_u >= 1 && g.n_u <= 4096, "row length must be in [1, 4096] (got %d)", g.n_u);
    w = {};
    w.u = plan->u;
    w.v = plan->v;
    w.ang = plan->angles;
    w.sdd = g.sdd;
    w.scan = g.total_scan_angle;
    w.short_scan = g.short_scan;
    w.curved = g.curved;
    for (int t = 0; t < bh_order; ++t) w.bh.c[t] = bh_coef[t];
    w.bh_order = bh_order;
    w.nviews = view_count;
    w.view0 = view_begin;
    return MCT_OK;
}

}  // namespace

extern "C" {

int mct_filter_rows(mct_ctx *ctx, const float *in, float *out, int64_t n_rows, int32_t n,
                    const float *taps, float scale, int32_t method) {
    MCT_CHECK(ctx && taps, "NULL argument");
    MCT_CHECK(n >= 1 && n <= 4096, "row length must be in [1, 4096] (got %d)", n);
    MCT_CHECK(n_rows >= 0, "negative row count");
    if (n_rows == 0) return MCT_OK;
    MCT_CHECK(in && out, "NULL argument");
    MCT_CHECK(in != out, "the row filter cannot run in place");
    if (method != MCT_FILTER_F32_DIRECT)
        return mct::filter_rows_mfma(ctx, in, out, n_rows, n, taps, scale, method, nullptr, 0);
    int rc = mct::upload_cached(ctx, &ctx->taps_dev, &ctx->taps_cap, &ctx->taps_hash, taps,
                                2 * (int64_t)n - 1);
    if (rc) return rc;
    size_t lds = sizeof(float) * (4 * (size_t)n + 2 * (size_t)n);
    int threads = (n + KB - 1) / KB;
    threads = ((threads + 63) / 64) * 64;
    threads = threads > 256 ? 256 : threads;
    int64_t blocks = (n_rows + RB - 1) / RB;
    MCT_CHECK(blocks < (int64_t)1 << 31, "too many rows (%lld)", (long long)n_rows);
    hipLaunchKernelGGL(ramp_direct_kernel, dim3((unsigned)blocks), dim3(threads), lds, ctx->stream,
                       in, out, n_rows, n, ctx->taps_dev, scale);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_weight_filter_rows(mct_plan *plan, const float *sgm, float *out, int32_t view_begin,
                           int32_t view_count, const float *bh_coef, int32_t bh_order,
                           const float *taps, float scale, int32_t method) {
    MCT_CHECK(plan && sgm && out && taps, "NULL argument");
    MCT_CHECK(sgm != out, "the row filter cannot run in place");
    mct::WeighSpec w;
    int rc = weigh_spec(plan, view_begin, view_count, bh_coef, bh_order, method, w);
    if (rc) return rc;
    const mct_geometry &g = plan->g;
    return mct::filter_rows_mfma(plan->ctx, sgm, out, (int64_t)g.n_v * view_count, g.n_u, taps, scale,
                                 method, &w, 0);
}

int mct_weight_filter_pack(mct_plan *plan, const float *sgm, void *packed, int32_t view_begin,
                           int32_t view_count, const float *bh_coef, int32_t bh_order,
                           const float *taps, float scale, int32_t method) {
    MCT_CHECK(plan && sgm && packed && taps, "NULL argument");
    mct::WeighSpec w;
    int rc = weigh_spec(plan, view_begin, view_count, bh_coef, bh_order, method, w);
    if (rc) return rc;
    const mct_geometry &g = plan->g;
    MCT_CHECK(g.n_v >= 2 && g.n_u >= 2 && (g.n_u - 1) % 64 != 0,
              "fused ROWS output needs at least 2 rows and n_u - 1 not a multiple of 64 "
              "(rows %d, n_u %d): filter, then mct_pack_sinogram", g.n_v, g.n_u);
    MCT_CHECK(g.n_v <= 65533, "ROWS packing takes at most 65533 rows");
    return mct::filter_rows_mfma(plan->ctx, sgm, reinterpret_cast<float *>(packed),
                                 (int64_t)g.n_v * view_count, g.n_u, taps, scale, method, &w, view_count);
}

}  // extern "C"
