// Sinogram weighting (BH polynomial, cosine, Parker), vertical Gaussian and
// the repack into the back-projector's tap layouts.
#include "mct_internal.h"
#include "mct_weigh.h"

namespace {

using mct::BHCoef;
using mct::kMaxBH;
using mct::weigh;

// Four consecutive elements per thread of a PLAIN [n_v][n_views][n_u]
// sinogram holding views [view0, view0 + nviews) of the scan (the Parker
// weight reads each one's angle).
__global__ void weight_kernel(float *__restrict__ sgm, int nv, int nviews, int nu,
                              const float *__restrict__ uarr, const float *__restrict__ varr,
                              const float *__restrict__ ang, float sdd, float scan,
                              int short_scan, int curved, BHCoef bh, int bh_order, int view0) {
    // grid: x over groups of 4 elements, y over views i, z over rows s
    const int j0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const int i = blockIdx.y, s = blockIdx.z;
    if (j0 >= nu) return;
    float *row = sgm + ((size_t)s * nviews + i) * nu;
    const bool vec = j0 + 4 <= nu && ((((uintptr_t)(row + j0)) & 15) == 0);
    float p[4];
    if (vec) {
        const float4 v = *reinterpret_cast<const float4 *>(row + j0);
        p[0] = v.x, p[1] = v.y, p[2] = v.z, p[3] = v.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) p[q] = j0 + q < nu ? row[j0 + q] : 0.f;
    }
    const float v = varr[s];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        p[q] = weigh(p[q], uarr[j0 + q < nu ? j0 + q : nu - 1], v, view0 + i, ang, sdd, scan,
                     short_scan, curved, bh, bh_order);
    if (vec) {
        *reinterpret_cast<float4 *>(row + j0) = make_float4(p[0], p[1], p[2], p[3]);
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (j0 + q < nu) row[j0 + q] = p[q];
    }
}

// run_mgfbp.py:1019-1031 (vertical Gaussian, +-10 rows, ascending n):
// out[i][jk] = det_h * sum_{|i-n|<=10} in[n][jk] * kern[i + rows-1 - n].
// A thread owns kVRows consecutive output rows of one (view, element) jk and
// holds the kVRows + 20 input rows they read in registers, so every input
// is read once per thread instead of once per output; each sum still runs
// over n in ascending order.
constexpr int kVRows = 16;
__global__ void vertical_kernel(const float *__restrict__ in, float *__restrict__ out, int rows,
                                int views, int nu, const float *__restrict__ kern, float det_h) {
    const int64_t plane = (int64_t)views * nu;
    const int64_t jk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jk >= plane) return;
    const int i0 = (int)blockIdx.y * kVRows;  // first output row of this thread
    float v[kVRows + 20];                      // input rows i0 - 10 .. i0 + kVRows + 9
#pragma unroll
    for (int t = 0; t < kVRows + 20; ++t) {
        const int n = i0 - 10 + t;
        v[t] = (n >= 0 && n < rows) ? in[n * plane + jk] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < kVRows; ++r) {
        const int i = i0 + r;
        if (i >= rows) break;
        float acc = 0.f;
#pragma unroll
        for (int d = -10; d <= 10; ++d) {
            const int n = i + d;
            if (n < 0 || n >= rows) continue;
            acc += v[r + 10 + d] * kern[i + (rows - 1) - n];
        }
        out[i * plane + jk] = acc * det_h;
    }
}

template <int DT> __device__ __forceinline__ uint16_t to_half_bits(float f) {
    return DT == MCT_F16 ? mct::f32_to_f16_rn(f) : mct::f32_to_bf16_rn(f);
}

// QUAD: out[j][r + 1][c] = bilinear coefficients of cell (r, c) of view j,
// r = -1 .. rows - 1: {a, b - a, c - a, d - c - b + a} with a = in[r][j][c],
// b = in[r][j][c+1], c = in[r+1][j][c], d = in[r+1][j][c+1] (each rounded
// once from FP64), rows -1 and `rows` read as zeros.  The two cells beyond
// the detector's rows let a back-projector interpolate any row coordinate in
// [-1, rows) without clamping it.  A thread walks kQuadRows cell rows of one
// column, so every detector row it reads serves two cells.
constexpr int kQuadRows = 16;
template <int DT>
__global__ void pack_quad_kernel(const float *__restrict__ in, int rows, int views, int nu,
                                 void *__restrict__ out) {
    // grid: x over detector cells c, y over views j, z over chunks of cell rows
    const int nc = nu - 1, ns = rows + 1;  // cell rows s = r + 1 in [0, rows]
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, sb = blockIdx.z * kQuadRows;
    if (c >= nc) return;
    const int se = sb + kQuadRows < ns ? sb + kQuadRows : ns;
    const int64_t rstride = (int64_t)views * nu;
    const float *p = in + (int64_t)j * nu + c;  // detector row 0
    // cell s = sr spans detector rows sr - 1 (a, b) and sr (c, d)
    double a = 0.0, b = 0.0;
    if (sb >= 1) {
        a = p[(sb - 1) * rstride];
        b = p[(sb - 1) * rstride + 1];
    }
    for (int sr = sb; sr < se; ++sr) {
        double cc = 0.0, d = 0.0;
        if (sr < rows) {
            cc = p[sr * rstride];
            d = p[sr * rstride + 1];
        }
        // bilinear coefficients: v(ru, rv) = q0 + ru q1 + rv (q2 + ru q3)
        const float q0 = (float)a, q1 = (float)(b - a), q2 = (float)(cc - a), q3 = (float)(d - cc - b + a);
        const size_t idx = ((size_t)j * ns + sr) * nc + c;
        if (DT == MCT_F32) {
            reinterpret_cast<float4 *>(out)[idx] = make_float4(q0, q1, q2, q3);
        } else {
            ushort4 h;
            h.x = to_half_bits<DT>(q0);
            h.y = to_half_bits<DT>(q1);
            h.z = to_half_bits<DT>(q2);
            h.w = to_half_bits<DT>(q3);
            reinterpret_cast<ushort4 *>(out)[idx] = h;
        }
        a = cc;
        b = d;
    }
}

// PAIR: out[r][j][c] = {in[r][j][c], in[r][j][c+1] - in[r][j][c]}
template <int DT>
__global__ void pack_pair_kernel(const float *__restrict__ in, int rows, int views, int nu,
                                 void *__restrict__ out) {
    const int nc = nu - 1;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y, j = blockIdx.z;
    if (c >= nc) return;
    const float *p = in + ((int64_t)r * views + j) * nu + c;
    const float q0 = p[0], q1 = (float)((double)p[1] - (double)p[0]);
    const size_t idx = ((size_t)r * views + j) * nc + c;
    if (DT == MCT_F32) {
        reinterpret_cast<float2 *>(out)[idx] = make_float2(q0, q1);
    } else {
        ushort2 h;
        h.x = to_half_bits<DT>(q0);
        h.y = to_half_bits<DT>(q1);
        reinterpret_cast<ushort2 *>(out)[idx] = h;
    }
}

// ROWS: out[j][r + 1][c] = {in[r][j][c], in[r][j][c+1] - in[r][j][c]} (the
// difference rounded once from FP64), r = -1 .. rows, rows -1 and `rows`
// zero: a view's rows contiguous, each with the PAIR taps of its elements.
// One element per thread: every load and store of a wavefront is contiguous
// (4 elements per thread with 16-byte loads measured 1.12 ms against 0.89
// ms at 256 x 1080 x 1024, its 8-byte stores strided 32 bytes apart).
template <int DT>
__global__ void pack_rows_kernel(const float *__restrict__ in, int rows, int views, int nu,
                                 void *__restrict__ out) {
    const int nc = nu - 1;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, s = blockIdx.z;  // stored row s = r + 1
    if (c >= nc) return;
    const int r = s - 1;
    float q0 = 0.f, q1 = 0.f;
    if (r >= 0 && r < rows) {
        const float *p = in + ((int64_t)r * views + j) * nu + c;
        q0 = p[0];
        q1 = (float)((double)p[1] - (double)p[0]);
    }
    const size_t idx = ((size_t)j * (rows + 2) + s) * nc + c;
    if (DT == MCT_F32) {
        reinterpret_cast<float2 *>(out)[idx] = make_float2(q0, q1);
    } else {
        ushort2 h;
        h.x = to_half_bits<DT>(q0);
        h.y = to_half_bits<DT>(q1);
        reinterpret_cast<ushort2 *>(out)[idx] = h;
    }
}

template <int DT>
__global__ void pack_plain_kernel(const float *__restrict__ in, int64_t n, void *__restrict__ out) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    if (DT == MCT_F32)
        reinterpret_cast<float *>(out)[idx] = in[idx];
    else
        reinterpret_cast<uint16_t *>(out)[idx] = to_half_bits<DT>(in[idx]);
}

}  // namespace

extern "C" {

int mct_weight_views(mct_plan *plan, float *sgm, int32_t view_begin, int32_t view_count,
                     const float *bh_coef, int32_t bh_order) {
    MCT_CHECK(plan && sgm, "NULL argument");
    MCT_CHECK(bh_order >= 0 && bh_order <= kMaxBH, "bh_order must be in [0, %d]", kMaxBH);
    MCT_CHECK(bh_order == 0 || bh_coef, "bh_coef is NULL");
    const mct_geometry &g = plan->g;
    MCT_CHECK(view_begin >= 0 && view_count >= 1 && view_count <= g.n_views - view_begin,
              "views [%d, %d + %d) outside the scan's %d views", view_begin, view_begin, view_count,
              g.n_views);
    mct_ctx *ctx = plan->ctx;
    BHCoef c = {};
    for (int t = 0; t < bh_order; ++t) c.c[t] = bh_coef[t];
    MCT_CHECK(g.n_v <= 65535 && view_count <= 65535, "rows and views must be <= 65535");
    hipLaunchKernelGGL(weight_kernel, dim3(mct::blocks_for((g.n_u + 3) / 4, 256), view_count, g.n_v), dim3(256), 0,
                       ctx->stream,
                       sgm, g.n_v, view_count, g.n_u, plan->u, plan->v, plan->angles, g.sdd,
                       g.total_scan_angle, g.short_scan, g.curved, c, bh_order, view_begin);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_weight_sinogram(mct_plan *plan, float *sgm, const float *bh_coef, int32_t bh_order) {
    MCT_CHECK(plan, "NULL argument");
    return mct_weight_views(plan, sgm, 0, plan->g.n_views, bh_coef, bh_order);
}

int mct_filter_vertical(mct_ctx *ctx, const float *in, float *out, int32_t rows, int32_t views,
                        int32_t n_u, const float *kern, float det_h) {
    MCT_CHECK(ctx && in && out && kern, "NULL argument");
    MCT_CHECK(in != out, "vertical filter cannot run in place");
    MCT_CHECK(rows > 0 && views > 0 && n_u > 0, "empty sinogram");
    int rc = mct::upload_cached(ctx, &ctx->vkern_dev, &ctx->vkern_cap, &ctx->vkern_hash, kern,
                                2 * (int64_t)rows - 1);
    if (rc) return rc;
    const int64_t plane = (int64_t)views * n_u;
    MCT_CHECK((rows + kVRows - 1) / kVRows <= 65535, "too many rows (%d)", rows);
    hipLaunchKernelGGL(vertical_kernel,
                       dim3(mct::blocks_for(plane, 256), (unsigned)((rows + kVRows - 1) / kVRows)),
                       dim3(256), 0, ctx->stream, in, out, rows, views, n_u, ctx->vkern_dev, det_h);
    MCT_LAUNCHED();
    return MCT_OK;
}

int64_t mct_packed_bytes(int32_t rows, int32_t views, int32_t n_u, int32_t dtype, int32_t layout) {
    if (rows <= 0 || views <= 0 || n_u <= 0) return -1;
    int64_t es = dtype == MCT_F32 ? 4 : 2;
    switch (layout) {
        case MCT_PLAIN: return (int64_t)rows * views * n_u * es;
        case MCT_QUAD:
            if (rows < 2 || n_u < 2) return -1;
            return (int64_t)(rows + 1) * views * (n_u - 1) * 4 * es;
        case MCT_PAIR:
            if (n_u < 2) return -1;
            return (int64_t)rows * views * (n_u - 1) * 2 * es;
        case MCT_ROWS:
            if (rows < 2 || n_u < 2) return -1;
            return (int64_t)(rows + 2) * views * (n_u - 1) * 2 * es;
        default: return -1;
    }
}

int mct_pack_sinogram(mct_ctx *ctx, const float *in, int32_t rows, int32_t views, int32_t n_u,
                      void *out, int32_t dtype, int32_t layout) {
    MCT_CHECK(ctx && in && out, "NULL argument");
    MCT_CHECK(dtype == MCT_F32 || dtype == MCT_F16 || dtype == MCT_BF16, "bad dtype %d", dtype);
    int64_t bytes = mct_packed_bytes(rows, views, n_u, dtype, layout);
    MCT_CHECK(bytes > 0, "bad packed shape (rows %d, views %d, n_u %d, layout %d)", rows, views,
              n_u, layout);
    MCT_CHECK(rows <= 65535 && views <= 65535, "rows and views must be <= 65535");
    MCT_CHECK(layout != MCT_ROWS || rows <= 65533, "ROWS packing takes at most 65533 rows");
    dim3 grid;
    if (layout == MCT_QUAD)
        grid = dim3(mct::blocks_for(n_u - 1, 256), views, (rows + 1 + kQuadRows - 1) / kQuadRows);
    else if (layout == MCT_PAIR)
        grid = dim3(mct::blocks_for(n_u - 1, 256), rows, views);
    else if (layout == MCT_ROWS)
        grid = dim3(mct::blocks_for(n_u - 1, 256), views, rows + 2);
    int64_t n = (int64_t)rows * views * n_u;
#define MCT_PACK_LAUNCH(K, G, ...)                                                            \
    do {                                                                                      \
        if (dtype == MCT_F32)                                                                 \
            hipLaunchKernelGGL(K<MCT_F32>, G, dim3(256), 0, ctx->stream, __VA_ARGS__);        \
        else if (dtype == MCT_F16)                                                            \
            hipLaunchKernelGGL(K<MCT_F16>, G, dim3(256), 0, ctx->stream, __VA_ARGS__);        \
        else                                                                                  \
            hipLaunchKernelGGL(K<MCT_BF16>, G, dim3(256), 0, ctx->stream, __VA_ARGS__);       \
    } while (0)
    if (layout == MCT_QUAD)
        MCT_PACK_LAUNCH(pack_quad_kernel, grid, in, rows, views, n_u, out);
    else if (layout == MCT_PAIR)
        MCT_PACK_LAUNCH(pack_pair_kernel, grid, in, rows, views, n_u, out);
    else if (layout == MCT_ROWS)
        MCT_PACK_LAUNCH(pack_rows_kernel, grid, in, rows, views, n_u, out);
    else
        MCT_PACK_LAUNCH(pack_plain_kernel, dim3(mct::blocks_for(n, 256)), in, n, out);
#undef MCT_PACK_LAUNCH
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // extern "C"
