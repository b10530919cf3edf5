// Helical sinogram rebinning.
//
// Reference: Mgfbp_helical.SinogramRebinning (run_mgfbp_helical.py:165-198),
// called once per reconstructed slice (run_mgfbp_helical.py:111-131).  For
// slice position z the rebinned fan sinogram takes view_count consecutive
// views starting at floor(k0 - view_count/2 + 1/2), where k0 is the view at
// which the detector centre row (after the vertical offset) crosses z; in
// each view the value is interpolated linearly between the two detector rows
// around v_pos = (z - k dz) * mag - offset (clamped to the first / last row)
// and scaled by sdd / sqrt(sdd^2 + v_pos^2).  Views past the scan (k < 0 or
// k > n_views - 2) give 0.
//
// Here every slice is rebinned in one launch: one thread per (slice, view,
// detector column), consecutive threads along the column, so a wavefront
// reads two contiguous row segments of one view.  FP32 in the reference's
// operation order.
#include "mct_internal.h"

namespace {

struct RebinArgs {
    int n_u, n_rows, n_views, view_count, nz;
    float off_v, det_h, sdd, mag, dz;
    const float *sgm;  // [n_rows][n_views][n_u]
    const float *z;    // [nz]
    float *out;        // [nz][view_count][n_u]
};

__global__ __launch_bounds__(256) void helical_rebin_kernel(RebinArgs a) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    const int zi = blockIdx.z;
    if (j >= a.n_u) return;
    const float z = a.z[zi];
    const float k0 = (z - a.off_v / a.mag) / a.dz;
    const float begin = floorf(k0 - 0.5f * (float)a.view_count + 0.5f);
    const float view = begin + (float)i;
    float val = 0.f;
    if (0.f <= view && view <= (float)(a.n_views - 2)) {
        const int k = (int)view;
        const float v_pos = (z - view * a.dz) * a.mag - a.off_v;
        const float v_idx = -v_pos / a.det_h + ((float)a.n_rows - 1.0f) / 2.0f;
        const size_t row = (size_t)a.n_views * a.n_u;
        const float *col = a.sgm + (size_t)k * a.n_u + j;
        if (v_idx >= (float)(a.n_rows - 1)) {
            val = col[(size_t)(a.n_rows - 1) * row];
        } else if (v_idx <= 0.f) {
            val = col[0];
        } else {
            const float fl = floorf(v_idx);
            const float w = v_idx - fl;
            const int r = (int)fl;
            val = col[(size_t)r * row] * (1.f - w) + col[(size_t)(r + 1) * row] * w;
        }
        val *= a.sdd / sqrtf(a.sdd * a.sdd + v_pos * v_pos);
    }
    a.out[((size_t)zi * a.view_count + i) * a.n_u + j] = val;
}

}  // namespace

extern "C" int mct_helical_rebin(mct_ctx *ctx, const mct_helical *h, const float *sgm,
                                 const float *z_pos, int32_t nz, float *out) {
    MCT_CHECK(ctx && h && sgm && z_pos && out, "NULL argument");
    MCT_CHECK(h->n_u >= 1 && h->n_rows >= 2 && h->n_views >= 2 && h->view_count >= 1,
              "bad sizes (n_u %d, rows %d, views %d, view_count %d)", h->n_u, h->n_rows, h->n_views,
              h->view_count);
    MCT_CHECK(h->view_count <= 65535, "view_count %d exceeds 65535", h->view_count);
    MCT_CHECK(nz >= 0 && nz <= 65535, "slice count %d outside [0, 65535]", nz);
    MCT_CHECK(h->det_h != 0.f && h->mag != 0.f && h->dis_per_view != 0.f,
              "det_h, mag and dis_per_view must be nonzero");
    if (nz == 0) return MCT_OK;
    MCT_HIP(hipSetDevice(ctx->device));
    int rc = mct::upload_cached(ctx, &ctx->aux_dev, &ctx->aux_cap, &ctx->aux_hash, z_pos, nz);
    if (rc) return rc;
    RebinArgs a;
    a.n_u = h->n_u;
    a.n_rows = h->n_rows;
    a.n_views = h->n_views;
    a.view_count = h->view_count;
    a.nz = nz;
    a.off_v = h->det_offset_vertical;
    a.det_h = h->det_h;
    a.sdd = h->sdd;
    a.mag = h->mag;
    a.dz = h->dis_per_view;
    a.sgm = sgm;
    a.z = ctx->aux_dev;
    a.out = out;
    hipLaunchKernelGGL(helical_rebin_kernel, dim3(mct::blocks_for(h->n_u, 256), h->view_count, nz),
                       dim3(256), 0, ctx->stream, a);
    MCT_LAUNCHED();
    return MCT_OK;
}
