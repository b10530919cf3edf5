// Ray-driven forward projection, ray mask and Poisson noise.
//
// Reference: Mgfpj.ForwardProjectionBilinear + BinSinogram
// (run_mgfpj.py:463-616), Mgfpj_v3.ForwardProjectionBilinear
// (run_mgfpj_v3.py:184-366), Mgfbp_ir.ForwardProjectionBilinear and
// GenEffectiveMapForwardProjectionAgent
// (run_mgfbp_iterative_recon_IRN_CGLS.py:494-833), Mgfpj.AddPossionNoise
// (run_mgfpj.py:643-654).
//
// The gantry is fixed and the object rotates: the source sits at
// (sid, 0, z_s), detector element (u, v) at (-(sdd - sid), u, v) (flat) and
// the ray is sampled at l = l_min + k*step*diag for k < count, each sample
// rotated by the view angle, tested against the volume (index and index+1
// inside) and interpolated tri/bilinearly; the sum of samples times
// step*diag is the line integral.  One thread traces one ray for every
// detector row and view in a single launch; the loop over k starts and ends
// at the ray's intersection with the volume's bounding box (FP64, widened by
// one step on each side), which skips samples the in-bounds test would
// reject anyway, so the sum is unchanged.
#include "mct_internal.h"

#include <cmath>
#include <vector>

namespace {

using mct::kBufferDword3;

struct FpArgs {
    int nx, nz, nu_os, nu_out, nrows, nviews, os, count;
    float pix, inv_pix, vh, inv_vh, x0, y0, zbase, cz;
    float sid, sdd, lmin, step, diag, zpv;
    int cone, curved, apply_pm;
    float c0, s0;  // cos / sin of angles[0] (P-matrix branch)
    const float *u, *v, *center_z;
    const float *pm_a, *pm_xs;  // P-matrix branch: A (9 per view), x_s (3 per view)
    const float2 *ft;  // {cos, sin} of the view angle
    const float *vol;
    const float *mask;  // NULL or [nrows][nviews][nu_out]
    float *out;         // [nrows][nviews][nu_out]
};

// A ray in the rotated (volume) frame: sample l sits at
// (ax + l bx, ay + l by, az + l bz).
struct Ray {
    float ax, ay, az, bx, by, bz;
};

// Ray of u index ui, row vi, view j.  Standard geometry (run_mgfpj.py:
// 531-551, 563-567): source (sid, 0, zshift), unit direction towards the
// detector element, rotated by the view angle.  P-matrix branch
// (run_mgfpj_v2.py:334-355): sample l sits at (l / dist) * A_j [ui, vi, 1]
// + x_s_j (z plus the helical shift), rotated by angles[0]; dist is the
// standard source-to-element distance, as in the reference.
__device__ __forceinline__ Ray make_ray(const FpArgs &a, int ui, int vi, int j) {
    float dx, dy;
    if (a.curved) {
        float gp = (-a.u[ui]) / a.sdd;
        dx = -a.sdd * cosf(gp) + a.sid;
        dy = -a.sdd * sinf(gp);
    } else {
        dx = -(a.sdd - a.sid);
        dy = a.u[ui];
    }
    const float zshift = a.zpv * (float)j;
    const float dz = a.v[vi] + zshift;
    const float ex = dx - a.sid, ey = dy - 0.0f, ez = dz - zshift;
    const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
    Ray r;
    if (!a.apply_pm) {
        const float ux = ex / dist, uy = ey / dist, uz = ez / dist;
        const float2 cs = a.ft[j];
        r.ax = a.sid * cs.x;
        r.ay = a.sid * cs.y;
        r.az = zshift;
        r.bx = ux * cs.x - uy * cs.y;
        r.by = uy * cs.x + ux * cs.y;
        r.bz = uz;
    } else {
        const float *A = a.pm_a + 9 * j;
        const float *xs = a.pm_xs + 3 * j;
        const float fu = (float)ui, fv = (float)vi;
        const float w0 = A[0] * fu + A[1] * fv + A[2];
        const float w1 = A[3] * fu + A[4] * fv + A[5];
        const float w2 = A[6] * fu + A[7] * fv + A[8];
        r.ax = xs[0] * a.c0 - xs[1] * a.s0;
        r.ay = xs[1] * a.c0 + xs[0] * a.s0;
        r.az = xs[2] + zshift;
        r.bx = (w0 * a.c0 - w1 * a.s0) / dist;
        r.by = (w1 * a.c0 + w0 * a.s0) / dist;
        r.bz = w2 / dist;
    }
    return r;
}

// Range [k_lo, k_hi) of sample indexes that can fall inside the volume's
// bounding box (plus one step of margin on each side); with cone = false
// only the in-plane box.
__device__ __forceinline__ void clip_range(const FpArgs &a, const Ray &r, float z0, bool cone,
                                           int &k_lo, int &k_hi) {
    double lo = -1e300, hi = 1e300;
    // box in rotated coordinates: x in [x0, x0+(nx-1)p], y in [y0-(nx-1)p, y0]
    double bx0 = a.x0, bx1 = (double)a.x0 + (double)(a.nx - 1) * a.pix;
    double by0 = (double)a.y0 - (double)(a.nx - 1) * a.pix, by1 = a.y0;
    auto slab = [&](double p, double d, double b0, double b1) {
        if (fabs(d) < 1e-300) {
            if (p < b0 || p > b1) { lo = 1e300; hi = -1e300; }
            return;
        }
        double t0 = (b0 - p) / d, t1 = (b1 - p) / d;
        if (t0 > t1) { double tt = t0; t0 = t1; t1 = tt; }
        lo = t0 > lo ? t0 : lo;
        hi = t1 < hi ? t1 : hi;
    };
    slab(r.ax, r.bx, bx0, bx1);
    slab(r.ay, r.by, by0, by1);
    if (cone) slab(r.az, r.bz, (double)z0, (double)z0 + (double)(a.nz - 1) * a.vh);
    if (!(lo <= hi)) {
        k_lo = k_hi = 0;
        return;
    }
    double dl = (double)a.step * (double)a.diag;
    double klo = floor((lo - a.lmin) / dl) - 1.0, khi = ceil((hi - a.lmin) / dl) + 2.0;
    klo = klo < 0.0 ? 0.0 : klo;
    khi = khi > (double)a.count ? (double)a.count : khi;
    k_lo = (int)klo;
    k_hi = (int)khi;
    if (k_hi < k_lo) k_hi = k_lo;
}

// Step range [v_lo, v_hi) that holds every sample of every ray of the view
// inside the volume's bounding box (same margins as clip_range): from the
// distance between the source (the rays' common origin (ax, ay, az)) and the
// box to the distance of its farthest corner.  It depends only on the view,
// so parts of it taken as segments split every ray of the view at the same
// steps, whatever rays share a wavefront.
__device__ __forceinline__ void view_range(const FpArgs &a, const Ray &r, float z0, bool cone,
                                           int &v_lo, int &v_hi) {
    const double bx0 = a.x0, bx1 = (double)a.x0 + (double)(a.nx - 1) * a.pix;
    const double by0 = (double)a.y0 - (double)(a.nx - 1) * a.pix, by1 = a.y0;
    const double bz0 = z0, bz1 = cone ? (double)z0 + (double)(a.nz - 1) * a.vh : (double)z0;
    const double sx = r.ax, sy = r.ay, sz = cone ? (double)r.az : bz0;
    auto near1 = [](double p, double b0, double b1) { return p < b0 ? b0 - p : (p > b1 ? p - b1 : 0.0); };
    auto far1 = [](double p, double b0, double b1) { return fmax(fabs(p - b0), fabs(p - b1)); };
    const double dn = sqrt(near1(sx, bx0, bx1) * near1(sx, bx0, bx1) + near1(sy, by0, by1) * near1(sy, by0, by1) +
                           near1(sz, bz0, bz1) * near1(sz, bz0, bz1));
    const double df = sqrt(far1(sx, bx0, bx1) * far1(sx, bx0, bx1) + far1(sy, by0, by1) * far1(sy, by0, by1) +
                           far1(sz, bz0, bz1) * far1(sz, bz0, bz1));
    // sample k sits at distance l_k from the source along a unit direction;
    // the P-matrix rays' parameter is not a distance, so they take every step
    const double dl = (double)a.step * (double)a.diag;
    double klo = floor((dn - a.lmin) / dl) - 1.0, khi = ceil((df - a.lmin) / dl) + 2.0;
    if (a.apply_pm) {
        klo = 0.0;
        khi = (double)a.count;
    }
    klo = klo < 0.0 ? 0.0 : klo;
    khi = khi > (double)a.count ? (double)a.count : khi;
    v_lo = (int)klo;
    v_hi = (int)khi;
    if (v_hi < v_lo) v_hi = v_lo;
}

// Line integral of one ray: run_mgfpj.py:554-605.
__device__ float trace(const FpArgs &a, int ui, int vi, int j) {
    const Ray r = make_ray(a, ui, vi, j);
    const float z0 = a.zbase + (a.center_z ? a.center_z[j] : a.cz);
    int k_lo, k_hi;
    clip_range(a, r, z0, a.cone, k_lo, k_hi);
    const float dl = a.step * a.diag;
    const int nx = a.nx;
    const size_t plane = (size_t)nx * nx;
    float sum = 0.f;
    for (int k = k_lo; k < k_hi; ++k) {
        float l = (float)k * a.step * a.diag + a.lmin;
        float xr = fmaf(l, r.bx, r.ax);
        float yr = fmaf(l, r.by, r.ay);
        float fx = (xr - a.x0) * a.inv_pix;
        float fy = (yr - a.y0) * (-a.inv_pix);
        float flx = floorf(fx), fly = floorf(fy);
        int xi = (int)flx, yi = (int)fly;
        if (xi < 0 || xi + 1 >= nx || yi < 0 || yi + 1 >= nx) continue;
        float xw = fx - flx, yw = fy - fly;
        if (a.cone) {
            float zr = fmaf(l, r.bz, r.az);
            float fz = (zr - z0) * a.inv_vh;
            float flz = floorf(fz);
            int zi = (int)flz;
            if (zi < 0 || zi + 1 >= a.nz) continue;
            float zw = fz - flz;
            const float *p = a.vol + (size_t)zi * plane + (size_t)yi * nx + xi;
            float v00 = p[0], v01 = p[1], v10 = p[nx], v11 = p[nx + 1];
            const float *q = p + plane;
            float w00 = q[0], w01 = q[1], w10 = q[nx], w11 = q[nx + 1];
            float lo0 = fmaf(xw, v01 - v00, v00), lo1 = fmaf(xw, v11 - v10, v10);
            float hi0 = fmaf(xw, w01 - w00, w00), hi1 = fmaf(xw, w11 - w10, w10);
            float lo = fmaf(yw, lo1 - lo0, lo0), hi = fmaf(yw, hi1 - hi0, hi0);
            sum += fmaf(zw, hi - lo, lo);
        } else {
            const float *p = a.vol + (size_t)vi * plane + (size_t)yi * nx + xi;
            float v00 = p[0], v01 = p[1], v10 = p[nx], v11 = p[nx + 1];
            float lo0 = fmaf(xw, v01 - v00, v00), lo1 = fmaf(xw, v11 - v10, v10);
            sum += fmaf(yw, lo1 - lo0, lo0);
        }
    }
    return sum * dl;
}

// One thread per output element (u_out, row, view): traces its os sub-rays
// and bins them (run_mgfpj.py:609-616).
__global__ __launch_bounds__(256) void fp_kernel(FpArgs a) {
    int uo = blockIdx.x * 64 + (threadIdx.x & 63);
    int vi = blockIdx.y * 4 + (threadIdx.x >> 6);
    int j = blockIdx.z;
    if (uo >= a.nu_out || vi >= a.nrows) return;
    size_t oidx = ((size_t)vi * a.nviews + j) * a.nu_out + uo;
    if (a.mask && a.mask[oidx] != 1.0f) {
        a.out[oidx] = 0.f;
        return;
    }
    float acc = 0.f;
    for (int i = 0; i < a.os; ++i) acc += trace(a, uo * a.os + i, vi, j);
    a.out[oidx] = acc / (float)a.os;
}

// ---------------------------------------------------------------------------
// Record forward projector.  What limits a one-thread-per-ray gather of the
// volume (measured with rocprofv3 counters at 512 x 512 x 256, 108 views of
// 1024 x 256): eight scalar taps per sample keep the vector L1 busy with
// little data, and 64 rays of neighbouring detector columns spread their
// samples along the view's detector direction, which for oblique views
// crosses a new volume row every cell: about 30 cache lines per wavefront
// load.
// Here the volume is first repacked into records of bilinear coefficients:
// record (y, x, z) = {a, b-a, c-a, (d-c)-(b-a)} of plane z with
// a = v[z][y][x], b = v[z][y][x+1], c = v[z][y+1][x], d = v[z][y+1][x+1]
// (FP32), x, y in [0, nx-2], stored z fastest: [y][x][z], 16 B each.  A
// cone-beam sample reads records z and z+1 (two 16-byte loads, adjacent); a
// fan-beam sample of detector row (= slice) r reads record r.  A wavefront
// traces a tile of 8 detector rows x 8 columns (cone; fan: the rows are
// slices, up to 64 of them per wavefront): at a given step the 8 rows of a
// column sit at neighbouring z (or are the same ray in neighbouring
// slices), so their records share cache lines whatever the view angle.  A
// lane whose sample falls in the cell its previous sample read skips the
// load and reuses those records (about half of the samples at step 0.2 of
// the voxel diagonal).  All lanes of a wavefront walk the union of their
// step ranges, so they sample at the same distance from the source; a
// sample outside a lane's own range or the volume reads past the end of the
// record buffer, which returns zeros, and adds +0.  Workgroups (four tiles
// stacked along the rows) run column tiles fastest, then views, then bands
// of rows, remapped so each XCD takes a contiguous part of that order: the
// views in flight on an XCD read one z-range of the records, which its L2
// and the last-level cache hold.  A record exists exactly when the
// reference's in-bounds test (index and index + 1 inside) passes.

// Bilinear coefficients of a cell, in FP32: v(xw, yw) = a + xw (b-a)
// + yw ((c-a) + xw ((d-c) - (b-a))).
__device__ __forceinline__ float4 cell_coef(float a, float b, float c, float d) {
    const float ba = b - a;
    return make_float4(a, ba, c - a, (d - c) - ba);
}

// Records of 16 cells along x of one row y and 16 planes per workgroup: the
// volume is read x fastest and the records written z fastest, through LDS;
// rec_y (optional) receives the same records in the order [x][y][z].
__global__ __launch_bounds__(256) void fp_pack_kernel(const float *__restrict__ vol, int nx, int nz,
                                                      float4 *__restrict__ rec,
                                                      float4 *__restrict__ rec_y) {
    __shared__ float row0[16][17], row1[16][17];  // [plane][x] of rows y and y + 1
    const int x0 = blockIdx.x * 16, y = blockIdx.y, z0 = blockIdx.z * 16;
    const int t = threadIdx.x;
    const size_t plane = (size_t)nx * nx;
    for (int i = t; i < 2 * 16 * 17; i += 256) {
        const int r = i / (16 * 17), e = i - r * (16 * 17), zz = e / 17, xx = e - zz * 17;
        const int gx = x0 + xx, gz = z0 + zz;
        const float v = (gx < nx && gz < nz) ? vol[(size_t)gz * plane + (size_t)(y + r) * nx + gx] : 0.f;
        if (r)
            row1[zz][xx] = v;
        else
            row0[zz][xx] = v;
    }
    __syncthreads();
    const int nc = nx - 1;
    const int xx = t >> 4, zz = t & 15, cx = x0 + xx, cz = z0 + zz;
    if (cx < nc && cz < nz) {
        const float4 c = cell_coef(row0[zz][xx], row0[zz][xx + 1], row1[zz][xx], row1[zz][xx + 1]);
        rec[((size_t)y * nc + cx) * nz + cz] = c;
        if (rec_y) rec_y[((size_t)cx * nc + y) * nz + cz] = c;
    }
}

// The volume with the slice index fastest: volt[(y * nx + x) * nz + z] =
// vol[z][y][x] (16 x positions of one row y and 16 slices per workgroup;
// read x fastest, written z fastest, through LDS).
__global__ __launch_bounds__(256) void fp_transpose_kernel(const float *__restrict__ vol, int nx, int nz,
                                                           float *__restrict__ volt) {
    __shared__ float tile[16][17];  // [slice][x]
    const int x0 = blockIdx.x * 16, y = blockIdx.y, z0 = blockIdx.z * 16;
    const int t = threadIdx.x;
    const size_t plane = (size_t)nx * nx;
    {
        const int zz = t >> 4, xx = t & 15, gx = x0 + xx, gz = z0 + zz;
        tile[zz][xx] = (gx < nx && gz < nz) ? vol[(size_t)gz * plane + (size_t)y * nx + gx] : 0.f;
    }
    __syncthreads();
    const int xx = t >> 4, zz = t & 15, gx = x0 + xx, gz = z0 + zz;
    if (gx < nx && gz < nz) volt[((size_t)y * nx + gx) * nz + gz] = tile[zz][xx];
}

__device__ __forceinline__ float cell_lerp(float4 q, float xw, float yw) {
    return fmaf(yw, fmaf(xw, q.w, q.z), fmaf(xw, q.y, q.x));
}

// (int)floor(x) in one instruction.  With fract(x), which equals x - floor(x)
// exactly for x >= 0 (the only samples that count), it replaces a floor, a
// convert and a subtraction per axis.
__device__ __forceinline__ int floor_to_int(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// c ? a : b per component (selects on values, never on addresses, so the
// records of a group stay in registers).
__device__ __forceinline__ float4 select4(bool c, float4 a, float4 b) {
    return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

__device__ __forceinline__ float4 load_rec(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const int o = __shfl_xor(v, m, 64);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const int o = __shfl_xor(v, m, 64);
        v = o > v ? o : v;
    }
    return v;
}

// A byte offset past every record: the buffer holds fewer than 2^32 - 64
// bytes, so this offset and the one 16 bytes after it both read zeros.
constexpr uint32_t kRecOut = 0u - 32u;

// Line integral of one ray from the records (the same samples, tests and
// order of summation as trace()).  [k_lo, k_hi) is the ray's clipped step
// range; with nseg > 1 the lane takes the steps of part seg of nseg of its
// view's step range (view_range: the same parts for every ray of the view,
// so the partial sums of a ray do not depend on the other rays of its
// wavefront).  Every lane of the wavefront must call this (active = false:
// a lane without a ray).  Samples go four at a time, loads first.
//
// The sample distance l of step k comes from ltab (the plan's table in the
// same FP32 expression order); k is wave-uniform, so it is a scalar load.
// (sxb, syb): byte strides of x and y in the record order rsrc covers.
template <bool CONE, bool REUSE>
__device__ float trace_rec(const FpArgs &a, const float *__restrict__ ltab, bool active, int ui,
                           int vi, int j, __amdgpu_buffer_rsrc_t rsrc, uint32_t sxb, uint32_t syb,
                           int seg, int nseg) {
    const Ray r = make_ray(a, active ? ui : 0, vi, j);
    const float z0 = a.zbase + (a.center_z ? a.center_z[j] : a.cz);
    int k_lo, k_hi;
    clip_range(a, r, z0, CONE, k_lo, k_hi);
    if (!active || k_hi <= k_lo) {
        k_lo = 0x7fffffff;
        k_hi = -1;
    }
    int w_lo = __builtin_amdgcn_readfirstlane(wave_min(k_lo));
    int w_hi = __builtin_amdgcn_readfirstlane(wave_max(k_hi));
    if (nseg > 1) {  // part seg of the view's step range
        int v_lo, v_hi;
        view_range(a, r, z0, CONE, v_lo, v_hi);
        const int64_t len = (int64_t)v_hi - v_lo;
        const int s0 = v_lo + (int)(len * seg / nseg), s1 = v_lo + (int)(len * (seg + 1) / nseg);
        w_lo = w_lo > s0 ? w_lo : s0;
        w_hi = w_hi < s1 ? w_hi : s1;
    }
    const float dl = a.step * a.diag;
    const uint32_t nc = (uint32_t)(a.nx - 1), nzr = (uint32_t)a.nz;
    const uint32_t slice = CONE ? 0u : (uint32_t)vi * 16u;
    float sum = 0.f;
    uint32_t prev = 0xffffffffu;  // offset of the records last loaded (none yet)
    float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
    constexpr int U = 4;
    for (int k0 = w_lo; k0 < w_hi; k0 += U) {
        uint32_t off[U];
        float xw[U], yw[U], zw[U];
#pragma unroll
        for (int i = 0; i < U; ++i) {
            const int k = k0 + i;
            const float l = ltab[k < a.count ? k : a.count - 1];
            const float xr = fmaf(l, r.bx, r.ax);
            const float yr = fmaf(l, r.by, r.ay);
            const float fx = (xr - a.x0) * a.inv_pix;
            const float fy = (yr - a.y0) * (-a.inv_pix);
            const int xi = floor_to_int(fx), yi = floor_to_int(fy);
            // k < w_hi: the last group may run past the wavefront's (or the
            // segment's) end.  A step outside the lane's own clipped range lies
            // outside the volume's box by at least a step, so the index test
            // rejects it.
            bool in = k < w_hi && (uint32_t)xi < nc && (uint32_t)yi < nc;
            xw[i] = __builtin_amdgcn_fractf(fx);
            yw[i] = __builtin_amdgcn_fractf(fy);
            uint32_t o = __umul24((uint32_t)yi, syb) + __umul24((uint32_t)xi, sxb);
            if (CONE) {
                const float zr = fmaf(l, r.bz, r.az);
                const float fz = (zr - z0) * a.inv_vh;
                const int zi = floor_to_int(fz);
                in = in && (uint32_t)zi < nzr - 1u;
                zw[i] = __builtin_amdgcn_fractf(fz);
                o += (uint32_t)zi * 16u;
            } else {
                zw[i] = 0.f;
                o += slice;
            }
            off[i] = in ? o : kRecOut;
        }
        // REUSE: a lane whose sample reads the records of its previous sample
        // points its load past the end of the buffer (range-checked: no cache
        // access) and takes the previous records instead, so every load of
        // the group is issued before any is waited for.  It pays where the
        // lanes of a wavefront share a ray (then they reuse together).  An
        // outside sample reads zeros (weights in [0, 1): +0).
        float4 q0[U], q1[U];
        bool same[U];
        uint32_t last = prev;
#pragma unroll
        for (int i = 0; i < U; ++i) {
            same[i] = REUSE && off[i] == last;
            last = off[i];
            const uint32_t o = same[i] ? kRecOut : off[i];
            q0[i] = load_rec(rsrc, o);
            if (CONE) q1[i] = load_rec(rsrc, o + 16u);
        }
        if (REUSE) {
#pragma unroll
            for (int i = 0; i < U; ++i) {
                q0[i] = select4(same[i], i ? q0[i - 1] : p0, q0[i]);
                if (CONE) q1[i] = select4(same[i], i ? q1[i - 1] : p1, q1[i]);
            }
            p0 = q0[U - 1];
            if (CONE) p1 = q1[U - 1];
            prev = off[U - 1];
        }
#pragma unroll
        for (int i = 0; i < U; ++i) {
            float val = cell_lerp(q0[i], xw[i], yw[i]);
            if (CONE) {
                const float hi = cell_lerp(q1[i], xw[i], yw[i]);
                val = fmaf(zw[i], hi - val, val);
            }
            sum += val;
        }
    }
    return sum * dl;
}

// fp_kernel on records (`bytes` of them, under 2^32 - 64).  A wavefront
// traces tr neighbouring detector rows x (64 / tr) neighbouring columns of
// one view (lanes row fastest).  wr = 4: a workgroup holds four such tiles
// stacked along the rows, and workgroups are numbered column block fastest,
// then view, then band of 4 tr rows; wr = 1 (rows too few for that): the
// wavefronts themselves are numbered column block fastest, then view, then
// row tile, four to a workgroup.  Each segment of a split ray range repeats
// the numbering; the whole is remapped so that each XCD takes a contiguous
// part (nwg: workgroups of one segment; the grid is padded to a multiple of
// the XCD count).
template <bool CONE, bool REUSE>
__device__ __forceinline__ void fp_rec_body(const FpArgs &a, const float *__restrict__ ltab,
                                            const float4 *rec, const float4 *rec_y, uint32_t bytes,
                                            int tr, int wr, int nub, int nwg, int nseg, int nxcd,
                                            float *partial) {
    const int g = mct::xcd_remap((int)blockIdx.x, (int)gridDim.x, nxcd);
    if (g >= nwg * nseg) return;  // padding
    const int seg = g / nwg, q = g - seg * nwg;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int tc = 64 / tr;
    int ub, j, rt;  // column block, view and row tile of this wavefront
    if (wr == 4) {  // the four tiles stacked along the rows
        const int t = q / nub;
        ub = q % nub;
        j = t % a.nviews;
        rt = (t / a.nviews) * 4 + wave;
    } else {  // wavefronts numbered one by one (column block fastest, then view, then row tile)
        const int w = q * 4 + wave, t = w / nub;
        ub = w % nub;
        j = t % a.nviews;
        rt = t / a.nviews;
    }
    if (j >= a.nviews) return;  // a whole wavefront past the last one (wr = 1)
    const int vi = rt * tr + (lane % tr);
    const int uo = ub * tc + lane / tr;
    const bool inside = uo < a.nu_out && vi < a.nrows;
    const size_t oidx = inside ? ((size_t)vi * a.nviews + j) * a.nu_out + uo : 0;
    const bool masked = inside && a.mask && a.mask[oidx] != 1.0f;
    const bool active = inside && !masked;
    // with a second order (few slices per tile, so a tile's lanes run along
    // the detector columns), the one whose fast axis follows the view's
    // detector direction (-sin, cos) (P-matrix: the rotated first column of A)
    bool ymaj = false;
    if (rec_y) {
        if (!a.apply_pm) {
            const float2 cs = a.ft[j];
            ymaj = fabsf(cs.x) >= fabsf(cs.y);
        } else {
            const float *A = a.pm_a + 9 * j;
            ymaj = fabsf(A[3] * a.c0 + A[0] * a.s0) >= fabsf(A[0] * a.c0 - A[3] * a.s0);
        }
    }
    const uint32_t colb = (uint32_t)a.nz * 16u, rowb = (uint32_t)(a.nx - 1) * colb;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)(ymaj ? rec_y : rec), 0, (int)bytes, kBufferDword3);
    const uint32_t sxb = ymaj ? rowb : colb, syb = ymaj ? colb : rowb;
    float acc = 0.f;
    for (int i = 0; i < a.os; ++i)
        acc += trace_rec<CONE, REUSE>(a, ltab, active, uo * a.os + i, inside ? vi : 0, j, rsrc, sxb,
                                      syb, seg, nseg);
    if (!inside) return;
    const float v = masked ? 0.f : acc / (float)a.os;
    if (nseg == 1)
        a.out[oidx] = v;
    else  // partial sum of segment seg; fp_segments_kernel adds them
        partial[(size_t)seg * a.nrows * a.nviews * a.nu_out + oidx] = v;
}

template <bool CONE, bool REUSE>
__global__ __launch_bounds__(256) void fp_rec_kernel(FpArgs a, const float *__restrict__ ltab,
                                                     const float4 *rec, const float4 *rec_y,
                                                     uint32_t bytes, int tr, int wr, int nub, int nwg,
                                                     int nseg, int nxcd, float *partial) {
    fp_rec_body<CONE, REUSE>(a, ltab, rec, rec_y, bytes, tr, wr, nub, nwg, nseg, nxcd, partial);
}

// Fan beam with at least 64 detector rows (slices): every row of a view and
// column is the same line (the rows differ only in the slice they sample),
// so a wavefront traces one ray in 64 slices (lane = slice) and computes the
// ray's sample positions, weights and in-bounds tests once, 64 steps at a
// time with a lane per step.  The steps a ray spends in one cell sample the
// same bilinear patch, so they add n a + Sx (b-a) + Sy (c-a) + Sxy ((d-c) -
// (b-a)) with n the step count and Sx, Sy, Sxy the sums of their weights
// wx, wy, wx wy: the first lane of each such run forms the sums (runs are a
// few steps long) and puts them in LDS, where every lane reads them
// (broadcast; one wavefront's LDS operations complete in order).  For each
// cell the ray enters, the four corner voxels of the 64 slices (256 B each,
// from the slice-fastest copy of the volume, which the last-level cache holds
// across views) are loaded once, four cells' loads in flight together, and
// turned into the cell's bilinear coefficients as the records are.  Same
// samples, in-bounds tests and coefficients as the record kernel; the
// per-cell grouping of the sum changes its rounding only.
__device__ __forceinline__ float4 corner_coef(float a, float b, float c, float d) {
    return cell_coef(a, b, c, d);
}

__device__ float trace_ray64(const FpArgs &a, const float *__restrict__ ltab,
                             __amdgpu_buffer_rsrc_t rsrc, uint32_t slice4, int ui, int j, int seg,
                             int nseg, float2 *__restrict__ wts, float4 *__restrict__ sums) {
    const Ray r = make_ray(a, ui, 0, j);
    int k_lo, k_hi;
    clip_range(a, r, 0.f, false, k_lo, k_hi);
    if (nseg > 1) {  // part seg of the view's step range
        int v_lo, v_hi;
        view_range(a, r, 0.f, false, v_lo, v_hi);
        const int64_t len = (int64_t)v_hi - v_lo;
        const int s0 = v_lo + (int)(len * seg / nseg), s1 = v_lo + (int)(len * (seg + 1) / nseg);
        k_lo = k_lo > s0 ? k_lo : s0;
        k_hi = k_hi < s1 ? k_hi : s1;
    }
    k_lo = __builtin_amdgcn_readfirstlane(k_lo);
    k_hi = __builtin_amdgcn_readfirstlane(k_hi);
    const int lane = threadIdx.x & 63;
    const uint32_t nc = (uint32_t)(a.nx - 1);
    const uint32_t xb = (uint32_t)a.nz * 4u, yb = (uint32_t)a.nx * xb;  // byte strides of x and y
    const float dl = a.step * a.diag;
    float sum = 0.f;
    // the cell of the last step of the previous chunk (inside or not) and its coefficients
    uint32_t c_off = kRecOut;
    int c_in = 0;
    float4 cur = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k0 = k_lo; k0 < k_hi; k0 += 64) {
        const int k = k0 + lane;
        const float l = ltab[k < a.count ? k : a.count - 1];
        const float xr = fmaf(l, r.bx, r.ax);
        const float yr = fmaf(l, r.by, r.ay);
        const float fx = (xr - a.x0) * a.inv_pix;
        const float fy = (yr - a.y0) * (-a.inv_pix);
        const int xi = floor_to_int(fx), yi = floor_to_int(fy);
        const bool in = k < k_hi && (uint32_t)xi < nc && (uint32_t)yi < nc;
        // this chunk's weights, read by the first lane of each run
        __builtin_amdgcn_wave_barrier();
        wts[lane] = make_float2(__builtin_amdgcn_fractf(fx), __builtin_amdgcn_fractf(fy));
        __builtin_amdgcn_wave_barrier();
        const uint32_t off = __umul24((uint32_t)yi, yb) + __umul24((uint32_t)xi, xb);
        uint32_t p_off = (uint32_t)__shfl_up((int)off, 1, 64);
        int p_in = __shfl_up((int)in, 1, 64);
        if (lane == 0) {
            p_off = c_off;
            p_in = c_in;
        }
        const bool fresh = in && !(p_in && p_off == off);
        const uint64_t m_in = __builtin_amdgcn_ballot_w64(in);
        const uint64_t m_new = __builtin_amdgcn_ballot_w64(fresh);
        c_off = (uint32_t)__builtin_amdgcn_readlane((int)off, 63);
        c_in = (m_in >> 63) & 1;
        // Each run of steps in one cell (a cell start, or step 0 continuing the
        // previous chunk's cell, up to the next cell start or step outside)
        // adds n a + Sx (b-a) + Sy (c-a) + Sxy ((d-c)-(b-a)) with the sums of
        // its steps' weights: the run's first lane forms them.
        const uint64_t m_stop = m_new | ~m_in;
        float sx = 0.f, sy = 0.f, sxy = 0.f;
        int e = lane;
        if (fresh || (lane == 0 && in)) {
            const uint64_t rest = lane < 63 ? m_stop >> (lane + 1) : 0;
            e = rest ? lane + 1 + (int)__builtin_ctzll(rest) : 64;
            for (int s2 = lane; s2 < e; ++s2) {
                const float2 w = wts[s2];
                sx += w.x;
                sy += w.y;
                sxy = fmaf(w.x, w.y, sxy);
            }
        }
        // the run sums of every lane, read back by all lanes (broadcast)
        __builtin_amdgcn_wave_barrier();
        sums[lane] = make_float4((float)(e - lane), sx, sy, sxy);
        __builtin_amdgcn_wave_barrier();
        if (!((m_new & 1) || !(m_in & 1))) {  // step 0 continues the previous chunk's cell
            const float4 c = sums[0];
            sum += fmaf(c.w, cur.w, fmaf(c.z, cur.z, fmaf(c.y, cur.y, c.x * cur.x)));
        }
        uint64_t m_rem = m_new;
        while (m_rem) {
            int n[4];
            float v[4][4];
            int cnt = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                n[t] = 0;
                if (m_rem) {
                    n[t] = __builtin_ctzll(m_rem);
                    m_rem &= m_rem - 1;
                    ++cnt;
                }
            }
            float4 c[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < cnt) {
                    // the cell's byte offset is uniform: the scalar offset of the loads
                    const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)off, n[t]);
                    const uint32_t o4 = o + slice4;
                    v[t][0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, o4, 0, 0));
                    v[t][1] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, o4 + xb, 0, 0));
                    v[t][2] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, o4 + yb, 0, 0));
                    v[t][3] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, o4 + yb + xb, 0, 0));
                    c[t] = sums[n[t]];
                }
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < cnt) {
                    const float4 q = corner_coef(v[t][0], v[t][1], v[t][2], v[t][3]);
                    sum += fmaf(c[t].w, q.w, fmaf(c[t].z, q.z, fmaf(c[t].y, q.y, c[t].x * q.x)));
                    cur = q;
                }
        }
    }
    return sum * dl;
}

// One wavefront per (output column, view, tile of 64 slices), numbered
// column fastest, then view, then slice tile, four to a workgroup; segments
// and XCD remapping as in fp_rec_kernel.
__global__ __launch_bounds__(256) void fp_ray64_kernel(FpArgs a, const float *__restrict__ ltab,
                                                       const float *volt, uint32_t bytes, int nwg,
                                                       int nseg, int nxcd, float *partial) {
    const int g = mct::xcd_remap((int)blockIdx.x, (int)gridDim.x, nxcd);
    if (g >= nwg * nseg) return;  // padding
    const int seg = g / nwg, q = g - seg * nwg;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int w = q * 4 + wave, t = w / a.nu_out;
    const int uo = w - t * a.nu_out, j = t % a.nviews, st = t / a.nviews;
    if (st * 64 >= a.nrows) return;  // a whole wavefront past the last tile
    const int vi = st * 64 + lane;
    const bool inside = vi < a.nrows;
    const size_t oidx = inside ? ((size_t)vi * a.nviews + j) * a.nu_out + uo : 0;
    const bool masked = inside && a.mask && a.mask[oidx] != 1.0f;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)volt, 0, (int)bytes, kBufferDword3);
    const uint32_t slice4 = (uint32_t)(inside ? vi : 0) * 4u;
    __shared__ float2 wts[4][64];   // per wavefront: the current chunk's step weights
    __shared__ float4 sums[4][64];  // per wavefront: the current chunk's run sums
    float acc = 0.f;
    for (int i = 0; i < a.os; ++i)
        acc += trace_ray64(a, ltab, rsrc, slice4, uo * a.os + i, j, seg, nseg, wts[wave], sums[wave]);
    if (!inside) return;
    const float v = masked ? 0.f : acc / (float)a.os;
    if (nseg == 1)
        a.out[oidx] = v;
    else  // partial sum of segment seg; fp_segments_kernel adds them
        partial[(size_t)seg * a.nrows * a.nviews * a.nu_out + oidx] = v;
}

// out[i] = sum of the nseg partial sums of ray i, in segment order.
__global__ void fp_segments_kernel(const float *__restrict__ partial, int64_t n, int nseg,
                                   float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = partial[i];
    for (int k = 1; k < nseg; ++k) s += partial[(size_t)k * n + i];
    out[i] = s;
}

// Mgfbp_ir.GenEffectiveMapForwardProjectionAgent (cone beam): 0 when a
// sample inside the image and the inscribed disc leaves the z range.
__global__ __launch_bounds__(256) void ray_mask_kernel(FpArgs a) {
    int ui = blockIdx.x * 64 + (threadIdx.x & 63);
    int vi = blockIdx.y * 4 + (threadIdx.x >> 6);
    int j = blockIdx.z;
    if (ui >= a.nu_out || vi >= a.nrows) return;
    size_t oidx = ((size_t)vi * a.nviews + j) * a.nu_out + ui;
    const Ray r = make_ray(a, ui, vi, j);
    const float z0 = a.zbase + (a.center_z ? a.center_z[j] : a.cz);
    int k_lo, k_hi;
    // clip in-plane only: the test fires when z leaves the range
    clip_range(a, r, z0, false, k_lo, k_hi);
    const float half = (float)a.nx / 2.0f;
    float m = 1.0f;
    for (int k = k_lo; k < k_hi; ++k) {
        float l = (float)k * a.step * a.diag + a.lmin;
        float xr = fmaf(l, r.bx, r.ax);
        float yr = fmaf(l, r.by, r.ay);
        int xi = (int)floorf((xr - a.x0) * a.inv_pix);
        int yi = (int)floorf((yr - a.y0) * (-a.inv_pix));
        if (xi < 0 || xi + 1 >= a.nx || yi < 0 || yi + 1 >= a.nx) continue;
        float dxi = (float)xi - half, dyi = (float)yi - half;
        if (dxi * dxi + dyi * dyi < half * half) {
            float zr = fmaf(l, r.bz, r.az);
            int zi = (int)floorf((zr - z0) * a.inv_vh);
            if (zi < 0 || zi + 1 >= a.nz) {
                m = 0.f;
                break;
            }
        }
    }
    a.out[oidx] = m;
}

__global__ void fill_kernel(float *p, int64_t n, float v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) + Box-Muller ----
__device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                             uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
}

__device__ __forceinline__ uint4 philox4x32_10(uint64_t ctr, uint64_t key) {
    uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32), c2 = 0, c3 = 0;
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    for (int r = 0; r < 10; ++r) {
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// Standard normal for element i: Box-Muller on the first two words of
// Philox(counter = i, key = seed), in FP64.
__device__ __forceinline__ double normal_of(uint64_t i, uint64_t seed) {
    uint4 r = philox4x32_10(i, seed);
    double u1 = ((double)r.x + 1.0) * (1.0 / 4294967296.0);  // (0, 1]
    double u2 = (double)r.y * (1.0 / 4294967296.0);          // [0, 1)
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

// run_mgfpj.py:643-654.
__global__ void noise_kernel(float *sgm, int64_t n, float photons, uint64_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t = photons * expf(-sgm[i]);
    t = t + (float)normal_of((uint64_t)i, seed) * sqrtf(t);
    if (t <= 0.f) t = 1e-6f;
    sgm[i] = logf(photons / t);
}

// Host side: scalars of run_mgfpj.py:486-527 in FP32, in the reference's order.
struct FpScalars {
    float x0, y0, zbase, lmin, diag;
    int count;
};

FpScalars fp_scalars(const mct_geometry &g, int count_mode) {
    FpScalars r;
    float nx = (float)g.img_dim, nz = (float)g.img_dim_z;
    float img_dimension = (float)g.img_dim * g.pix;
    float img_dimension_z = (float)g.img_dim_z * g.vox_h;
    r.x0 = -(nx - 1.0f) / 2.0f * g.pix + g.cx;
    r.y0 = -(nx - 1.0f) / 2.0f * (-g.pix) + g.cy;
    r.zbase = -(nz - 1.0f) / 2.0f * g.vox_h;
    float D = std::sqrt(2.0f * (img_dimension * img_dimension) + img_dimension_z * img_dimension_z);
    r.lmin = g.sid - D / 2.0f;
    float lmax = g.sid + D / 2.0f;
    r.diag = std::sqrt(2.0f * (g.pix * g.pix) + g.vox_h * g.vox_h);
    float span = count_mode == 1 ? D : (lmax - r.lmin);
    r.count = (int)std::floor(span / (g.fpj_step * r.diag));
    return r;
}

}  // namespace

namespace mct {

// Distance the gantry moves per view for a helical scan (run_mgfpj.py:518-523).
float helical_z_per_view(const mct_plan *p, const float *angles_host) {
    const mct_geometry &g = p->g;
    if (!g.helical) return 0.f;
    int nv = g.n_views;
    float tsa = std::fabs(angles_host[nv - 1] - angles_host[0]) / (float)(nv - 1) * (float)nv;
    float laps = tsa / (kPI * 2.f);
    return g.helical_pitch * (laps / (float)nv) * (p->adv * (float)g.n_v_full) / (g.sdd / g.sid);
}

}  // namespace mct

namespace {

int fill_args(mct_plan *p, FpArgs &a, int count_mode) {
    const mct_geometry &g = p->g;
    FpScalars s = fp_scalars(g, count_mode);
    a = FpArgs{};
    a.nx = g.img_dim;
    a.nz = g.img_dim_z;
    a.nrows = g.n_v;
    a.nviews = g.n_views;
    a.count = s.count;
    a.pix = g.pix;
    a.inv_pix = 1.0f / g.pix;
    a.vh = g.vox_h;
    a.inv_vh = g.vox_h > 0.f ? 1.0f / g.vox_h : 0.f;
    a.x0 = s.x0;
    a.y0 = s.y0;
    a.zbase = s.zbase;
    a.cz = g.cz;
    a.sid = g.sid;
    a.sdd = g.sdd;
    a.lmin = s.lmin;
    a.step = g.fpj_step;
    a.diag = s.diag;
    a.cone = g.cone;
    a.curved = g.curved;
    a.u = p->u;
    a.v = p->v;
    a.center_z = p->center_z;
    a.ft = p->fp_view;
    a.zpv = p->zpv;
    a.apply_pm = g.apply_pmatrix;
    a.c0 = p->cos_a0;
    a.s0 = p->sin_a0;
    a.pm_a = p->matrix_a;
    a.pm_xs = p->x_s;
    MCT_CHECK(!g.apply_pmatrix || (p->matrix_a && p->x_s),
              "P-matrix forward projection needs matrix_a and x_s (9 and 3 floats per view)");
    return MCT_OK;
}

// The plan's table of sample distances l[k] = k * step * diag + l_min for
// k < count, in the FP32 expression order of the direct kernel's loop
// (rebuilt when the scalars change; one synchronisation when built).
int ensure_step_table(mct_plan *p, const FpArgs &a) {
    if (a.count <= 0) return MCT_OK;
    if (p->fp_ltab && p->fp_ltab_n == a.count && p->fp_ltab_key[0] == a.step &&
        p->fp_ltab_key[1] == a.diag && p->fp_ltab_key[2] == a.lmin)
        return MCT_OK;
    mct_ctx *ctx = p->ctx;
    if (p->fp_ltab) {
        MCT_HIP(hipStreamSynchronize(ctx->stream));
        MCT_HIP(hipFree(p->fp_ltab));
        p->fp_ltab = nullptr;
        p->fp_ltab_n = 0;
    }
    std::vector<float> l((size_t)a.count);
    for (int k = 0; k < a.count; ++k) l[(size_t)k] = (float)k * a.step * a.diag + a.lmin;
    MCT_HIP(hipMalloc((void **)&p->fp_ltab, sizeof(float) * l.size()));
    MCT_HIP(hipMemcpyAsync(p->fp_ltab, l.data(), sizeof(float) * l.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    MCT_HIP(hipStreamSynchronize(ctx->stream));  // l is released on return
    p->fp_ltab_n = a.count;
    p->fp_ltab_key[0] = a.step;
    p->fp_ltab_key[1] = a.diag;
    p->fp_ltab_key[2] = a.lmin;
    return MCT_OK;
}

}  // namespace

extern "C" {

int mct_forward_project(mct_plan *plan, const float *vol, float *out, const float *ray_mask,
                        int32_t oversample) {
    MCT_CHECK(plan && vol && out, "NULL argument");
    const mct_geometry &g = plan->g;
    MCT_CHECK(oversample >= 1, "oversample must be >= 1 (got %d)", oversample);
    MCT_CHECK(g.n_u % oversample == 0, "n_u (%d) is not a multiple of oversample (%d)", g.n_u,
              oversample);
    MCT_CHECK(!ray_mask || oversample == 1, "a ray mask requires oversample 1");
    MCT_CHECK(g.fpj_step > 0.f, "fpj_step must be positive");
    MCT_CHECK(!g.cone || g.vox_h > 0.f, "cone beam needs vox_h > 0");
    MCT_CHECK(g.cone || g.n_v <= g.img_dim_z, "fan beam: row %d has no slice (%d slices)", g.n_v - 1,
              g.img_dim_z);
    FpArgs a;
    int rc0 = fill_args(plan, a, plan->fp_count_mode);
    if (rc0) return rc0;
    a.os = oversample;
    a.nu_os = g.n_u;
    a.nu_out = g.n_u / oversample;
    a.vol = vol;
    a.mask = ray_mask;
    a.out = out;
    mct_ctx *ctx = plan->ctx;
    if (g.img_dim < 2) {
        // a one-pixel image has no cell: no sample passes the in-bounds test
        // 0 <= index, index + 1 < dim (run_mgfpj.py:572), so every line sum is 0
        const int64_t nout = (int64_t)a.nrows * a.nviews * a.nu_out;
        hipLaunchKernelGGL(fill_kernel, dim3(mct::blocks_for(nout, 256)), dim3(256), 0, ctx->stream, out,
                           nout, 0.f);
        MCT_LAUNCHED();
        ctx->fp_last = ctx->fp_kernel == MCT_FP_DIRECT ? MCT_FP_DIRECT : MCT_FP_CELLS;
        return MCT_OK;
    }
    dim3 grid((a.nu_out + 63) / 64, (a.nrows + 3) / 4, a.nviews);
    if (ctx->fp_kernel != MCT_FP_DIRECT) {
        const int64_t nx = g.img_dim, nz = g.img_dim_z;
        const int64_t nc = nx - 1;
        const int64_t bytes = nc * nc * nz * 16;  // records [y][x][z]
        // tile rows: cone 8 (x 8 columns); fan up to 64 slices of one ray
        int tr = 8;
        if (!g.cone) {
            tr = 1;
            while (tr < 64 && tr < a.nrows) tr *= 2;
        }
        const int tc = 64 / tr;
        const int wr = (int64_t)4 * tr <= a.nrows ? 4 : 1;
        const int64_t nub = (a.nu_out + tc - 1) / tc;
        const int64_t rtiles = ((int64_t)a.nrows + tr - 1) / tr;
        const int64_t nwg = wr == 4 ? nub * a.nviews * ((rtiles + 3) / 4)
                                    : (nub * a.nviews * rtiles + 3) / 4;
        // Small problems (e.g. one fan slice) would not fill the device with
        // one lane per ray: split each ray's steps over nseg <= 16 lanes
        // (toward 256 wavefronts per CU in all, so the last round of
        // wavefronts is a small part of the launch) and add the partial sums
        // in a second pass.  (512^2 fan slice, 240 views: 0.46 ms with 32
        // wavefronts per CU, 0.35 ms with 256.)  The split counts the rays
        // (wavefronts of 64 rays), so a problem of n rows with the target
        // scaled by n splits each ray exactly as a one-row problem does (the
        // count is rows x the wavefronts of one row, so no rounding couples
        // the rows); the parts come from each ray's own step range.
        const int64_t target = (int64_t)ctx->cu_count * (ctx->fp_split ? ctx->fp_split : 256);
        const int64_t nwaves_rays = ((int64_t)a.nu_out * a.nviews + 63) / 64 * a.nrows;
        int64_t nseg = (target + nwaves_rays - 1) / nwaves_rays;
        nseg = nseg < 1 ? 1 : (nseg > 16 ? 16 : nseg);
        const int nxcd = ctx->xcd_count > 0 ? ctx->xcd_count : 1;
        const int64_t nblocks = (nwg * nseg + nxcd - 1) / nxcd * nxcd;
        const int64_t nout = (int64_t)a.nrows * a.nviews * a.nu_out;
        // Fan beam, at least 64 rows, no P-matrix: a wavefront per ray in 64
        // slices, on the slice-fastest copy of the volume (32-bit byte
        // offsets, 24-bit row stride).
        const int64_t tbytes = nx * nx * nz * 4;
        if (!g.cone && !g.apply_pmatrix && a.nrows >= 64 && tbytes < ((int64_t)1 << 32) - 64 &&
            nx * nz * 4 < ((int64_t)1 << 24) && nx <= 65535) {
            const int64_t nw = (int64_t)a.nu_out * a.nviews * (((int64_t)a.nrows + 63) / 64);
            const int64_t nwg64 = (nw + 3) / 4;
            int64_t nseg64 = (target + nw - 1) / nw;
            nseg64 = nseg64 < 1 ? 1 : (nseg64 > 16 ? 16 : nseg64);
            const int64_t nb64 = (nwg64 * nseg64 + nxcd - 1) / nxcd * nxcd;
            if (nb64 < ((int64_t)1 << 31)) {
                int rc = ensure_step_table(plan, a);
                if (rc) return rc;
                rc = mct::ensure_buffer(ctx, &ctx->fp_pack, &ctx->fp_pack_cap, tbytes / 4);
                if (rc) return rc;
                if (nseg64 > 1) {
                    rc = mct::ensure_buffer(ctx, &ctx->fp_partial, &ctx->fp_partial_cap, nseg64 * nout);
                    if (rc) return rc;
                }
                const dim3 tgrid((unsigned)((nx + 15) / 16), (unsigned)nx, (unsigned)((nz + 15) / 16));
                hipLaunchKernelGGL(fp_transpose_kernel, tgrid, dim3(256), 0, ctx->stream, vol, g.img_dim,
                                   g.img_dim_z, ctx->fp_pack);
                hipLaunchKernelGGL(fp_ray64_kernel, dim3((unsigned)nb64), dim3(256), 0, ctx->stream, a,
                                   plan->fp_ltab, ctx->fp_pack, (uint32_t)tbytes, (int)nwg64, (int)nseg64,
                                   nxcd, ctx->fp_partial);
                if (nseg64 > 1)
                    hipLaunchKernelGGL(fp_segments_kernel, dim3(mct::blocks_for(nout, 256)), dim3(256), 0,
                                       ctx->stream, ctx->fp_partial, nout, (int)nseg64, out);
                MCT_LAUNCHED();
                ctx->fp_last = MCT_FP_CELLS;
                return MCT_OK;
            }
        }
        // 32-bit record offsets (and 24-bit row strides for the address
        // arithmetic); gridDim.y of the pack launch
        const bool fits = nc >= 1 && nz >= 1 && bytes < ((int64_t)1 << 32) - 64 &&
                          nc * nz * 16 < ((int64_t)1 << 24) && nc * nc * nz * 16 < ((int64_t)1 << 32) &&
                          nblocks < ((int64_t)1 << 31) && nc <= 65535;
        if (fits) {
            int rc = ensure_step_table(plan, a);
            if (rc) return rc;
            const bool two = !g.cone && tr <= 4;  // a second, y-major order
            rc = mct::ensure_buffer(ctx, &ctx->fp_pack, &ctx->fp_pack_cap, (two ? 2 : 1) * (bytes / 4));
            if (rc) return rc;
            if (nseg > 1) {
                rc = mct::ensure_buffer(ctx, &ctx->fp_partial, &ctx->fp_partial_cap, nseg * nout);
                if (rc) return rc;
            }
            float4 *rec = reinterpret_cast<float4 *>(ctx->fp_pack);
            float4 *rec_y = two ? reinterpret_cast<float4 *>(ctx->fp_pack + bytes / 4) : nullptr;
            const dim3 pgrid((unsigned)((nc + 15) / 16), (unsigned)nc, (unsigned)((nz + 15) / 16));
            hipLaunchKernelGGL(fp_pack_kernel, pgrid, dim3(256), 0, ctx->stream, vol, g.img_dim,
                               g.img_dim_z, rec, rec_y);
#define FP_REC_LAUNCH(CONE, REUSE)                                                                  \
    hipLaunchKernelGGL((fp_rec_kernel<CONE, REUSE>), dim3((unsigned)nblocks), dim3(256), 0, ctx->stream, \
                       a, plan->fp_ltab, rec, rec_y, (uint32_t)bytes, tr, wr, (int)nub, (int)nwg,      \
                       (int)nseg, nxcd, ctx->fp_partial)
            if (g.cone)
                FP_REC_LAUNCH(true, true);
            else if (tr == 64)  // the lanes of a wavefront are one ray in 64 slices
                FP_REC_LAUNCH(false, true);
            else
                FP_REC_LAUNCH(false, false);
#undef FP_REC_LAUNCH
            if (nseg > 1)
                hipLaunchKernelGGL(fp_segments_kernel, dim3(mct::blocks_for(nout, 256)), dim3(256), 0,
                                   ctx->stream, ctx->fp_partial, nout, (int)nseg, out);
            MCT_LAUNCHED();
            ctx->fp_last = MCT_FP_CELLS;
            return MCT_OK;
        }
        if (ctx->fp_kernel == MCT_FP_CELLS)
            return mct::set_error(MCT_ERR_UNSUPPORTED,
                                  "record forward projection: the records (%lld B) do not fit "
                                  "32-bit offsets",
                                  (long long)bytes);
    }
    hipLaunchKernelGGL(fp_kernel, grid, dim3(256), 0, ctx->stream, a);
    MCT_LAUNCHED();
    ctx->fp_last = MCT_FP_DIRECT;
    return MCT_OK;
}

int mct_ray_mask(mct_plan *plan, float *mask) {
    MCT_CHECK(plan && mask, "NULL argument");
    const mct_geometry &g = plan->g;
    mct_ctx *ctx = plan->ctx;
    int64_t n = (int64_t)g.n_v * g.n_views * g.n_u;
    if (!g.cone) {
        hipLaunchKernelGGL(fill_kernel, dim3(mct::blocks_for(n, 256)), dim3(256), 0, ctx->stream, mask,
                           n, 1.0f);
        MCT_LAUNCHED();
        return MCT_OK;
    }
    MCT_CHECK(g.vox_h > 0.f, "cone beam needs vox_h > 0");
    FpArgs a;
    int rc0 = fill_args(plan, a, plan->fp_count_mode);
    if (rc0) return rc0;
    a.os = 1;
    a.nu_os = g.n_u;
    a.nu_out = g.n_u;
    a.out = mask;
    dim3 grid((a.nu_out + 63) / 64, (a.nrows + 3) / 4, a.nviews);
    hipLaunchKernelGGL(ray_mask_kernel, grid, dim3(256), 0, ctx->stream, a);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_add_noise(mct_ctx *ctx, float *sgm, int64_t n, float photon_number, uint64_t seed) {
    MCT_CHECK(ctx && sgm, "NULL argument");
    MCT_CHECK(photon_number > 0.f, "photon number must be positive");
    if (n <= 0) return MCT_OK;
    hipLaunchKernelGGL(noise_kernel, dim3(mct::blocks_for(n, 256)), dim3(256), 0, ctx->stream, sgm, n,
                       photon_number, seed);
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // extern "C"
