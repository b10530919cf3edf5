// Vector kernels of the IRN-CGLS / TV / PICCS loop with FP64 scalars on the
// device.
//
// Reference: Mgfbp_ir.TikhonovSol, GenerateWR, TVPenaltyVal, Dt_W_D,
// LossValCalc (run_mgfbp_iterative_recon_IRN_CGLS.py:184-286), where they are
// float32 NumPy on the host.  Here the reductions are FP64 and deterministic:
// a fixed grid of kReduceBlocks workgroups walks the array in a fixed order,
// then one workgroup adds the partials in a fixed tree, so a result does not
// depend on what else runs on the device.
#include "mct_internal.h"

#include <mutex>

namespace {

constexpr int kReduceBlocks = 1024;
constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// partials[b] -> out[0] (one workgroup).
__global__ __launch_bounds__(kThreads) void finish_sum_kernel(const double *__restrict__ partials,
                                                              int nb, double *__restrict__ out) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += kThreads) v += partials[i];
    v = block_sum(v);
    if (threadIdx.x == 0) out[0] = v;
}

__global__ __launch_bounds__(kThreads) void dot_kernel(const float *__restrict__ a,
                                                       const float *__restrict__ b, int64_t n,
                                                       double *__restrict__ partials) {
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)kReduceBlocks * kThreads)
        v += (double)a[i] * (double)b[i];
    v = block_sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

__global__ __launch_bounds__(kThreads) void sq_diff_kernel(const float *__restrict__ a,
                                                           const float *__restrict__ b,
                                                           const float *__restrict__ mb, int64_t n,
                                                           double *__restrict__ partials) {
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)kReduceBlocks * kThreads) {
        double bb = mb ? (double)b[i] * (double)mb[i] : (double)b[i];
        double d = (double)a[i] - bb;
        v += d * d;
    }
    v = block_sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// Forward differences of y = x - prior at (z, yy, xx): ux along axis 1, uy
// along axis 2, uz along axis 0 (only when cz, i.e. slices are coupled and
// nz > 2); 0 at the last index.
struct Grad {
    double ux, uy, uz;
};
__device__ __forceinline__ double val(const float *x, const float *prior, int64_t i) {
    return prior ? (double)x[i] - (double)prior[i] : (double)x[i];
}
__device__ __forceinline__ Grad grad_at(const float *x, const float *prior, int nz, int ny, int nx,
                                        int cz, int z, int yy, int xx) {
    int64_t i = ((int64_t)z * ny + yy) * nx + xx;
    double c = val(x, prior, i);
    Grad g;
    g.ux = yy + 1 < ny ? val(x, prior, i + nx) - c : 0.0;
    g.uy = xx + 1 < nx ? val(x, prior, i + 1) - c : 0.0;
    g.uz = (cz && z + 1 < nz) ? val(x, prior, i + (int64_t)ny * nx) - c : 0.0;
    return g;
}

__global__ __launch_bounds__(kThreads) void tv_weight_kernel(const float *__restrict__ x,
                                                             const float *__restrict__ prior, int nz,
                                                             int ny, int nx, int cz, double beta,
                                                             float *__restrict__ wr) {
    int64_t n = (int64_t)nz * ny * nx;
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int xx = (int)(i % nx), yy = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    Grad g = grad_at(x, prior, nz, ny, nx, cz, z, yy, xx);
    wr[i] = (float)(1.0 / sqrt(g.ux * g.ux + g.uy * g.uy + g.uz * g.uz + beta));
}

__global__ __launch_bounds__(kThreads) void tv_value_kernel(const float *__restrict__ x,
                                                            const float *__restrict__ prior, int nz,
                                                            int ny, int nx, int cz, double beta,
                                                            double *__restrict__ partials) {
    int64_t n = (int64_t)nz * ny * nx;
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)kReduceBlocks * kThreads) {
        int xx = (int)(i % nx), yy = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        Grad g = grad_at(x, prior, nz, ny, nx, cz, z, yy, xx);
        v += sqrt(g.ux * g.ux + g.uy * g.uy + g.uz * g.uz + beta);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// Dt_W_D (run_mgfbp_iterative_recon_IRN_CGLS.py:267-286): with the weighted
// differences gy = W*ux, gx = W*uy, gz = W*uz,
//   out = -(gy[y] - gy[y-1]) [1 <= y <= ny-2] - (gx[x] - gx[x-1]) [1 <= x <= nx-2]
//         - (gz[z] - gz[z-1]) [1 <= z <= nz-2, cz].
__global__ __launch_bounds__(kThreads) void dtwd_kernel(const float *__restrict__ x,
                                                        const float *__restrict__ wr, int nz, int ny,
                                                        int nx, int cz, float scale,
                                                        const float *__restrict__ base,
                                                        float *__restrict__ out) {
    int64_t n = (int64_t)nz * ny * nx;
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int xx = (int)(i % nx), yy = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    const int64_t sy = nx, sz = (int64_t)ny * nx;
    double acc = 0.0;
    if (yy >= 1 && yy <= ny - 2) {
        double g1 = (double)wr[i] * ((double)x[i + sy] - (double)x[i]);
        double g0 = (double)wr[i - sy] * ((double)x[i] - (double)x[i - sy]);
        acc -= g1 - g0;
    }
    if (xx >= 1 && xx <= nx - 2) {
        double g1 = (double)wr[i] * ((double)x[i + 1] - (double)x[i]);
        double g0 = (double)wr[i - 1] * ((double)x[i] - (double)x[i - 1]);
        acc -= g1 - g0;
    }
    if (cz && z >= 1 && z <= nz - 2) {
        double g1 = (double)wr[i] * ((double)x[i + sz] - (double)x[i]);
        double g0 = (double)wr[i - sz] * ((double)x[i] - (double)x[i - sz]);
        acc -= g1 - g0;
    }
    double b = base ? (double)base[i] : 0.0;
    out[i] = (float)(b + (double)scale * acc);
}

__global__ __launch_bounds__(kThreads) void axpby_kernel(float a, const float *__restrict__ x, float b,
                                                         const float *__restrict__ y,
                                                         float *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}

// alpha = s[0]/s[1]; x += alpha d; r -= alpha q; partial sums of <r,r>.
__global__ __launch_bounds__(kThreads) void cg_update_kernel(float *__restrict__ x,
                                                             float *__restrict__ r,
                                                             const float *__restrict__ d,
                                                             const float *__restrict__ q,
                                                             const double *__restrict__ s, int64_t n,
                                                             double *__restrict__ partials) {
    const double alpha = s[0] / s[1];
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)kReduceBlocks * kThreads) {
        x[i] = (float)((double)x[i] + alpha * (double)d[i]);
        float rn = (float)((double)r[i] - alpha * (double)q[i]);
        r[i] = rn;
        v += (double)rn * (double)rn;
    }
    v = block_sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// beta = s[2]/s[0]; d = r + beta d.  s[0] is rewritten by a later kernel.
__global__ __launch_bounds__(kThreads) void cg_direction_kernel(float *__restrict__ d,
                                                                const float *__restrict__ r,
                                                                const double *__restrict__ s,
                                                                int64_t n) {
    const double beta = s[2] / s[0];
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) d[i] = (float)((double)r[i] + beta * (double)d[i]);
}

__global__ void copy_scalar_kernel(double *s) { s[0] = s[2]; }

int finish(mct_ctx *ctx, double *out) {
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, ctx->partials,
                       kReduceBlocks, out);
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // namespace

// Scratch of mct_dot_f64_noalloc: the kReduceBlocks FP64 partial sums of one
// reduction (8 KiB), the same for every length, so no input is too large for
// it.  Each device holds its own copy for the life of the process; host code
// uses the device address taken once per device by noalloc_scratch_init,
// never the symbol itself.  Every call writes all the partials before the
// finishing kernel reads them, so nothing has to start at zero.
static __device__ __align__(256) double scratch[kReduceBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace mct {

// Once per device, with that device current (mct_create calls it).
int noalloc_scratch_init(int device) {
    MCT_CHECK(device >= 0 && device < kMaxDevices, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return MCT_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(MCT_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                         hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
    return MCT_OK;
}

}  // namespace mct

extern "C" {

int mct_dot_f64(mct_ctx *ctx, const float *a, const float *b, int64_t n, double *out) {
    MCT_CHECK(ctx && a && b && out, "NULL argument");
    MCT_CHECK(n >= 0, "negative length");
    int rc = mct::ensure_partials(ctx, kReduceBlocks);
    if (rc) return rc;
    hipLaunchKernelGGL(dot_kernel, dim3(kReduceBlocks), dim3(kThreads), 0, ctx->stream, a, b, n,
                       ctx->partials);
    MCT_LAUNCHED();
    return finish(ctx, out);
}

int mct_dot_f64_noalloc(mct_ctx *ctx, const float *a, const float *b, int64_t n, double *out) {
    MCT_CHECK(ctx && a && b && out, "NULL argument");
    MCT_CHECK(n >= 0, "negative length");
    MCT_CHECK(ctx->device >= 0 && ctx->device < kMaxDevices, "device %d beyond %d", ctx->device,
              kMaxDevices);
    double *part;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        part = g_scratch_ptr[ctx->device];
    }
    MCT_CHECK(part, "no-allocation scratch is not set up on device %d", ctx->device);
    hipLaunchKernelGGL(dot_kernel, dim3(kReduceBlocks), dim3(kThreads), 0, ctx->stream, a, b, n, part);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, part, kReduceBlocks,
                       out);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_sq_diff_f64(mct_ctx *ctx, const float *a, const float *b, const float *mb, int64_t n,
                    double *out) {
    MCT_CHECK(ctx && a && b && out, "NULL argument");
    MCT_CHECK(n >= 0, "negative length");
    int rc = mct::ensure_partials(ctx, kReduceBlocks);
    if (rc) return rc;
    hipLaunchKernelGGL(sq_diff_kernel, dim3(kReduceBlocks), dim3(kThreads), 0, ctx->stream, a, b, mb,
                       n, ctx->partials);
    MCT_LAUNCHED();
    return finish(ctx, out);
}

int mct_tv_weight(mct_ctx *ctx, const float *x, const float *prior, int32_t nz, int32_t ny,
                  int32_t nx, int32_t couple_z, double beta, float *wr) {
    MCT_CHECK(ctx && x && wr, "NULL argument");
    MCT_CHECK(nz > 0 && ny > 0 && nx > 0, "empty volume");
    int64_t n = (int64_t)nz * ny * nx;
    hipLaunchKernelGGL(tv_weight_kernel, dim3(mct::blocks_for(n, kThreads)), dim3(kThreads), 0,
                       ctx->stream, x, prior, nz, ny, nx, (int)(couple_z && nz > 2), beta, wr);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_tv_value(mct_ctx *ctx, const float *x, const float *prior, int32_t nz, int32_t ny,
                 int32_t nx, int32_t couple_z, double beta, double *out) {
    MCT_CHECK(ctx && x && out, "NULL argument");
    MCT_CHECK(nz > 0 && ny > 0 && nx > 0, "empty volume");
    int rc = mct::ensure_partials(ctx, kReduceBlocks);
    if (rc) return rc;
    hipLaunchKernelGGL(tv_value_kernel, dim3(kReduceBlocks), dim3(kThreads), 0, ctx->stream, x, prior,
                       nz, ny, nx, (int)(couple_z && nz > 2), beta, ctx->partials);
    MCT_LAUNCHED();
    return finish(ctx, out);
}

int mct_dtwd(mct_ctx *ctx, const float *x, const float *wr, int32_t nz, int32_t ny, int32_t nx,
             int32_t couple_z, float scale, const float *base, float *out) {
    MCT_CHECK(ctx && x && wr && out, "NULL argument");
    MCT_CHECK(nz > 0 && ny > 0 && nx > 0, "empty volume");
    MCT_CHECK(out != x && out != wr, "out must not alias x or wr");
    int64_t n = (int64_t)nz * ny * nx;
    hipLaunchKernelGGL(dtwd_kernel, dim3(mct::blocks_for(n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       x, wr, nz, ny, nx, (int)(couple_z && nz > 2), scale, base, out);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_axpby(mct_ctx *ctx, float a, const float *x, float b, const float *y, float *out, int64_t n) {
    MCT_CHECK(ctx && x && y && out, "NULL argument");
    if (n <= 0) return MCT_OK;
    hipLaunchKernelGGL(axpby_kernel, dim3(mct::blocks_for(n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       a, x, b, y, out, n);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_cg_update(mct_ctx *ctx, float *x, float *r, const float *d, const float *q, double *s,
                  int64_t n) {
    MCT_CHECK(ctx && x && r && d && q && s, "NULL argument");
    MCT_CHECK(n >= 0, "negative length");
    int rc = mct::ensure_partials(ctx, kReduceBlocks);
    if (rc) return rc;
    hipLaunchKernelGGL(cg_update_kernel, dim3(kReduceBlocks), dim3(kThreads), 0, ctx->stream, x, r, d,
                       q, s, n, ctx->partials);
    MCT_LAUNCHED();
    return finish(ctx, s + 2);
}

int mct_cg_direction(mct_ctx *ctx, float *d, const float *r, double *s, int64_t n) {
    MCT_CHECK(ctx && d && r && s, "NULL argument");
    if (n > 0) {
        hipLaunchKernelGGL(cg_direction_kernel, dim3(mct::blocks_for(n, kThreads)), dim3(kThreads), 0,
                           ctx->stream, d, r, s, n);
        MCT_LAUNCHED();
    }
    hipLaunchKernelGGL(copy_scalar_kernel, dim3(1), dim3(1), 0, ctx->stream, s);
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // extern "C"
