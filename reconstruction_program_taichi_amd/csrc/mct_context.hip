// Context, error reporting and geometry plans.
#include "mct_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

namespace {
thread_local char g_last_error[512] = "";

template <typename T>
int upload(T **dst, const T *src, size_t n) {
    *dst = nullptr;
    if (!src || n == 0) return MCT_OK;
    MCT_HIP(hipMalloc((void **)dst, sizeof(T) * n));
    MCT_HIP(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return MCT_OK;
}

void free_plan_arrays(mct_plan *p) {
    void *ptrs[] = {p->angles,   p->u,        p->v,       p->pmatrix, p->center_z, p->matrix_a,
                    p->x_s,      p->bp_view,  p->fp_view, p->bp_iv,   p->bp_win,   p->fp_ltab};
    for (void *q : ptrs)
        if (q) hipFree(q);
}

// FP32 cos/sin of an FP32 angle: evaluated in double and rounded once, so the
// table is reproducible on any host.
inline float cos_rn(float a) { return (float)std::cos((double)a); }
inline float sin_rn(float a) { return (float)std::sin((double)a); }

// Order-preserving map between floats and integers (for bisection over
// every representable float).
inline int64_t fkey(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i >= 0 ? (int64_t)i : -(int64_t)(i & 0x7fffffff);
}
inline float funkey(int64_t k) {
    uint32_t u = k >= 0 ? (uint32_t)k : ((uint32_t)(-k) | 0x80000000u);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Smallest float in (bad, good] where pred holds, pred monotone along the
// float order with pred(bad) false and pred(good) true (good > bad), or the
// largest float in [good, bad) when good > bad.
template <class P>
float boundary(P pred, float good, float bad) {
    int64_t g = fkey(good), b = fkey(bad);
    while (g - b > 1 || b - g > 1) {
        int64_t m = g + (b - g) / 2;
        if (pred(funkey(m)))
            g = m;
        else
            b = m;
    }
    return funkey(g);
}

// The reference's range tests in FP32, in its expression order.
struct URange {
    float u0, du;
    int nu;
    bool operator()(float u) const {
        float uidx = (u - u0) / du;
        return !(uidx < 0.f || uidx + 1.f > (float)(nu - 1));
    }
};
struct VRange {
    float v0, det_h, adv, dv;
    int nv;
    bool operator()(float d) const {
        float vidx = d / det_h * adv / dv;
        float vf = std::floor(vidx);
        return !(vf < 0.f || vf + 1.f > (float)(nv - 1));
    }
};
}  // namespace

namespace mct {

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int refuse_frozen(const char *what, int64_t have, int64_t need) {
    return set_error(MCT_ERR_INVALID,
                     "context is frozen (a captured graph may use its scratch): %s would change "
                     "(%lld -> %lld)", what, (long long)have, (long long)need);
}

int ensure_partials(mct_ctx *ctx, int64_t n) {
    if (ctx->partials_cap >= n) return MCT_OK;
    if (ctx->frozen) return refuse_frozen("reduction partials", ctx->partials_cap, n);
    if (ctx->partials) {
        MCT_HIP(hipStreamSynchronize(ctx->stream));
        MCT_HIP(hipFree(ctx->partials));
    }
    ctx->partials = nullptr;
    ctx->partials_cap = 0;
    int64_t cap = n < 8192 ? 8192 : n;
    MCT_HIP(hipMalloc((void **)&ctx->partials, sizeof(double) * cap));
    ctx->partials_cap = cap;
    return MCT_OK;
}

int ensure_buffer(mct_ctx *ctx, float **buf, int64_t *cap, int64_t n) {
    if (*cap >= n) return MCT_OK;
    if (ctx->frozen) return refuse_frozen("a scratch buffer (floats)", *cap, n);
    if (*buf) {
        MCT_HIP(hipStreamSynchronize(ctx->stream));
        MCT_HIP(hipFree(*buf));
    }
    *buf = nullptr;
    *cap = 0;
    int64_t c = n < 4096 ? 4096 : n;
    MCT_HIP(hipMalloc((void **)buf, sizeof(float) * c));
    *cap = c;
    return MCT_OK;
}

int upload_cached(mct_ctx *ctx, float **buf, int64_t *cap, uint64_t *hash, const float *host,
                  int64_t n) {
    const uint64_t h = fnv1a(host, sizeof(float) * (size_t)n, (uint64_t)n);
    if (*buf && *cap >= n && *hash == h) return MCT_OK;
    if (ctx->frozen) return refuse_frozen("a cached device table (floats)", *cap, n);
    int rc = ensure_buffer(ctx, buf, cap, n);
    if (rc) return rc;
    *hash = 0;  // invalid until the copy is queued
    MCT_HIP(hipMemcpyAsync(*buf, host, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    *hash = h;
    return MCT_OK;
}

float div_factor_of(const mct_geometry &g) {
    // run_mgfbp.py:1072-1084, FP32 as in the Taichi kernel.
    const float PI = kPI;
    float ascan = std::fabs(g.total_scan_angle);
    float num_rounds = std::floor(ascan / (PI * 2.f));
    float div = 1.f;
    if (g.short_scan) {
        float remain = ascan - num_rounds * 2.f * PI;
        if (ascan < 2.f * PI)
            div = 1.f;
        else if (remain < PI)
            div = 1.f / (num_rounds * 2.f);
        else if (PI < remain && remain < 2.f * PI)
            div = 1.f / (num_rounds * 2.f + 1.f);
    } else {
        div = 1.f / (num_rounds * 2.f);
    }
    return div;
}

}  // namespace mct

extern "C" {

const char *mct_last_error(void) { return g_last_error; }

int mct_abi_version(void) { return MCT_ABI_VERSION; }

int mct_create(int device, mct_ctx **out) {
    if (!out) return mct::set_error(MCT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    MCT_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return mct::set_error(MCT_ERR_INVALID, "device %d out of range (%d devices)", device, count);
    MCT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    MCT_HIP(hipGetDeviceProperties(&prop, device));
    const int rc = mct::noalloc_scratch_init(device);
    if (rc) return rc;
    mct_ctx *c = new mct_ctx();
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    c->xcd_count = (strstr(prop.gcnArchName, "gfx950") || strstr(prop.gcnArchName, "gfx942")) ? 8 : 1;
    *out = c;
    return MCT_OK;
}

int mct_destroy(mct_ctx *ctx) {
    if (!ctx) return MCT_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (ctx->partials) hipFree(ctx->partials);
    if (ctx->toep) hipFree(ctx->toep);
    if (ctx->fsplit) hipFree(ctx->fsplit);
    if (ctx->taps_dev) hipFree(ctx->taps_dev);
    if (ctx->vkern_dev) hipFree(ctx->vkern_dev);
    if (ctx->aux_dev) hipFree(ctx->aux_dev);
    if (ctx->fp_pack) hipFree(ctx->fp_pack);
    if (ctx->fp_partial) hipFree(ctx->fp_partial);
    if (ctx->bp_rec) hipFree(ctx->bp_rec);
    for (mct::BatchSlot &b : ctx->batch_slot) {
        if (b.ev) hipEventSynchronize(b.ev);
        if (b.host) hipHostFree(b.host);
        if (b.dev) hipFree(b.dev);
        if (b.ev) hipEventDestroy(b.ev);
    }
    delete ctx;
    return MCT_OK;
}

int mct_set_stream(mct_ctx *ctx, void *stream) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    ctx->stream = (hipStream_t)stream;
    return MCT_OK;
}

int mct_synchronize(mct_ctx *ctx) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    MCT_HIP(hipStreamSynchronize(ctx->stream));
    return MCT_OK;
}

int mct_copy_views(mct_ctx *ctx, void *stream, const float *host_sgm, int32_t n_v, int32_t n_views,
                   int32_t n_u, int32_t view_begin, int32_t view_count, float *dev_block) {
    MCT_CHECK(ctx && host_sgm && dev_block, "NULL argument");
    MCT_CHECK(n_v > 0 && n_u > 0, "empty sinogram");
    MCT_CHECK(view_begin >= 0 && view_count >= 1 && view_count <= n_views - view_begin,
              "views [%d, %d + %d) outside the sinogram's %d views", view_begin, view_begin,
              view_count, n_views);
    const size_t width = (size_t)view_count * n_u * sizeof(float);
    const size_t pitch = (size_t)n_views * n_u * sizeof(float);
    MCT_HIP(hipMemcpy2DAsync(dev_block, width, host_sgm + (size_t)view_begin * n_u, pitch, width,
                             (size_t)n_v, hipMemcpyHostToDevice,
                             stream ? (hipStream_t)stream : ctx->stream));
    return MCT_OK;
}

int mct_set_bp_tile(mct_ctx *ctx, int32_t zc, int32_t batch) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    if (!(zc == 0 || zc == 4 || zc == 8 || zc == 16 || zc == 32))
        return mct::set_error(MCT_ERR_INVALID, "zc must be 0, 4, 8, 16 or 32 (got %d)", zc);
    if (!(batch == 0 || batch == 4 || batch == 8 || batch == 16))
        return mct::set_error(MCT_ERR_INVALID, "batch must be 0, 4, 8 or 16 (got %d)", batch);
    ctx->bp_zc = zc;
    ctx->bp_batch = batch;
    return MCT_OK;
}

int mct_set_bp_kernel(mct_ctx *ctx, int32_t kind) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    if (kind < MCT_BP_AUTO || kind > MCT_BP_LDS)
        return mct::set_error(MCT_ERR_INVALID, "back-projector kind must be 0, 1 or 2 (got %d)", kind);
    ctx->bp_kernel = kind;
    return MCT_OK;
}

int mct_bp_last_kernel(const mct_ctx *ctx) { return ctx ? ctx->bp_last : -1; }

int mct_set_fp_kernel(mct_ctx *ctx, int32_t kind) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    if (kind < MCT_FP_AUTO || kind > MCT_FP_CELLS)
        return mct::set_error(MCT_ERR_INVALID, "forward-projector kind must be 0, 1 or 2 (got %d)", kind);
    ctx->fp_kernel = kind;
    return MCT_OK;
}

int mct_fp_last_kernel(const mct_ctx *ctx) { return ctx ? ctx->fp_last : -1; }

int mct_set_fp_split(mct_ctx *ctx, int32_t waves_per_cu) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    if (waves_per_cu < 0 || waves_per_cu > 4096)
        return mct::set_error(MCT_ERR_INVALID, "waves_per_cu must be in [0, 4096] (got %d)", waves_per_cu);
    ctx->fp_split = waves_per_cu;
    return MCT_OK;
}

int mct_set_frozen(mct_ctx *ctx, int32_t frozen) {
    if (!ctx) return mct::set_error(MCT_ERR_INVALID, "ctx is NULL");
    ctx->frozen = frozen ? 1 : 0;
    return MCT_OK;
}

int mct_plan_create(mct_ctx *ctx, const mct_geometry *g, mct_plan **out) {
    if (!out) return mct::set_error(MCT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    MCT_CHECK(ctx != nullptr, "ctx is NULL");
    MCT_CHECK(g != nullptr, "geometry is NULL");
    MCT_CHECK(g->img_dim > 0 && g->img_dim_z > 0, "image dimensions must be positive (%d, %d)",
              g->img_dim, g->img_dim_z);
    MCT_CHECK(g->n_views >= 2, "at least two views are required (got %d)", g->n_views);
    MCT_CHECK(g->n_u >= 2, "at least two detector elements per row are required (got %d)", g->n_u);
    MCT_CHECK(g->n_v >= 1, "at least one detector row is required (got %d)", g->n_v);
    MCT_CHECK(g->angles && g->u && g->v, "angles, u and v arrays must be given");
    MCT_CHECK(!g->cone || g->n_v >= 2, "cone-beam geometry needs at least two detector rows");
    MCT_CHECK(g->view_mode >= 1 && g->view_mode <= 3, "view_mode must be 1, 2 or 3 (got %d)",
              g->view_mode);
    MCT_CHECK(!g->apply_pmatrix || g->pmatrix, "apply_pmatrix set without a pmatrix array");
    MCT_CHECK(g->pix > 0.f, "pixel size must be positive");
    MCT_CHECK(g->u[1] != g->u[0], "u[1] == u[0]");
    MCT_CHECK(!g->cone || g->v[1] != g->v[0], "v[1] == v[0]");
    MCT_HIP(hipSetDevice(ctx->device));

    mct_plan *p = new mct_plan();
    p->ctx = ctx;
    p->device = ctx->device;
    p->g = *g;
    p->g.angles = p->g.u = p->g.v = p->g.pmatrix = p->g.center_z = p->g.matrix_a = p->g.x_s = nullptr;
    const int nv = g->n_views;
    int rc = MCT_OK;
    if (!rc) rc = upload(&p->angles, g->angles, nv);
    if (!rc) rc = upload(&p->u, g->u, g->n_u);
    if (!rc) rc = upload(&p->v, g->v, g->n_v);
    if (!rc) rc = upload(&p->pmatrix, g->pmatrix, g->pmatrix ? 12 * (size_t)nv : 0);
    if (!rc) rc = upload(&p->center_z, g->center_z, g->center_z ? (size_t)nv : 0);
    if (!rc) rc = upload(&p->matrix_a, g->matrix_a, g->matrix_a ? 9 * (size_t)nv : 0);
    if (!rc) rc = upload(&p->x_s, g->x_s, g->x_s ? 3 * (size_t)nv : 0);
    if (!rc) {
        // run_mgfbp.py:1105-1111 (dbeta) and :1122 (angle without rotation).
        std::vector<float4> bt(nv);
        std::vector<float2> ft(nv);
        const float *a = g->angles;
        const float divf = mct::div_factor_of(*g);
        for (int j = 0; j < nv; ++j) {
            float ar = a[j] - g->img_rot;
            float d = (j == nv - 1) ? std::fabs(a[nv - 1] - a[0]) / (float)(nv - 1) : std::fabs(a[j + 1] - a[j]);
            bt[j] = make_float4(cos_rn(ar), sin_rn(ar), d, g->sid * d * divf);
            ft[j] = make_float2(cos_rn(a[j]), sin_rn(a[j]));
        }
        rc = upload(&p->bp_view, bt.data(), nv);
        if (!rc) rc = upload(&p->fp_view, ft.data(), nv);
    }
    if (rc) {
        free_plan_arrays(p);
        delete p;
        return rc;
    }
    p->u0 = g->u[0];
    p->du = g->u[1] - g->u[0];
    p->v0 = g->v[0];
    if (g->n_v >= 2) {
        p->dv = g->v[1] - g->v[0];
        p->adv = std::fabs(p->dv);
    }
    p->cos_rot = cos_rn(g->img_rot);
    p->sin_rot = sin_rn(g->img_rot);
    p->cos_a0 = cos_rn(g->angles[0]);
    p->sin_a0 = sin_rn(g->angles[0]);
    p->div_factor = mct::div_factor_of(*g);
    p->fp_count_mode = g->fpj_count_mode;
    p->zpv = mct::helical_z_per_view(p, g->angles);
    {
        URange ur{p->u0, p->du, g->n_u};
        float umid = p->u0 + p->du * (float)(g->n_u - 1) * 0.5f;
        if (ur(umid)) {
            p->ua = boundary(ur, umid, -3e38f);
            p->ub = boundary(ur, umid, 3e38f);
        } else {  // no u passes (n_u == 2 with a degenerate spacing): empty interval
            p->ua = 1.f;
            p->ub = -1.f;
        }
        if (g->cone && g->n_v >= 2 && g->det_h > 0.f) {
            VRange vr{p->v0, g->det_h, p->adv, p->dv, g->n_v};
            float dmid = (float)(g->n_v - 1) * 0.5f * g->det_h * (p->dv > 0.f ? 1.f : -1.f);
            if (vr(dmid)) {
                float e0 = boundary(vr, dmid, -3e38f), e1 = boundary(vr, dmid, 3e38f);
                p->da = e0 < e1 ? e0 : e1;
                p->db = e0 < e1 ? e1 : e0;
            } else {
                p->da = 1.f;
                p->db = -1.f;
            }
        }
    }
    *out = p;
    return MCT_OK;
}

int mct_plan_destroy(mct_plan *p) {
    if (!p) return MCT_OK;
    // The plan's context may already be destroyed (callers may release the
    // two in either order), so nothing here reads p->ctx; hipFree waits for
    // the device's outstanding work before it releases the memory.
    hipSetDevice(p->device);
    free_plan_arrays(p);
    delete p;
    return MCT_OK;
}

}  // extern "C"
