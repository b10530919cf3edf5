// Pixel-driven back-projection kernels.
//
// Reference: Mgfbp.BackProjectionPixelDriven (run_mgfbp.py:1062-1187) and
// Mgfbp_ir.BackProjectionPixelDrivenPerView + SetTruncatedRegionToZero
// (run_mgfbp_iterative_recon_IRN_CGLS.py:372-481).
//
// Axial fast path (no P-matrix): one thread owns an (x, y) column and ZC
// consecutive slices of it; a wavefront covers 8 x 8 columns.  Everything
// that depends only on (column, view) - the source distance L, the
// magnification, the detector u position, the view weight and the range of
// slices whose projection stays on the detector - is computed once per view,
// so the slice loop is branch-free: per voxel-view update one FMA for the
// detector row coordinate, a clamp (direct kernel only), a convert, a fract,
// one address multiply-add, the packed taps and four FMAs.  All ZC loads of a
// view are independent and in flight together.
//
// The filtered sinogram arrives pre-packed, so the taps of one update are one
// or two vector loads:
//   ROWS [view][row+2][col-1] x {a, b-a} for cone beam (element rows -1 ..
//     rows, rows beyond the detector zero): a cell's four coefficients come
//     from two vertically neighbouring entries;
//   QUAD [view][row+1][col-1] x {a, b-a, c-a, d-c-b+a} for cone beam (cell
//     rows -1 .. rows-1): one 16-byte (FP32) or 8-byte (FP16/BF16) load;
//   PAIR [row][view][col-1] x {a, b-a} for fan beam.
// Workgroups are remapped so that each XCD owns a contiguous range of slice
// chunks: its L2 then only holds the detector rows that range projects onto.
//
// Mask semantics (run_mgfbp.py:1142-1144, 1170-1173): a voxel whose
// projection leaves the detector in any view is 0.  The geometry plan turns
// the reference's FP32 range tests into exact intervals on u = fl(mag * yr)
// and d = fl(fl(mag * z) - v[0]) (both tests are monotone in those values),
// so the kernels decide the mask exactly as the reference's FP32 arithmetic
// does (the library is built with -ffp-contract=off; mag = sdd / L is an
// IEEE division).  Because d is monotone in the slice index, the slices that
// pass the v test in a view form an interval; it is located with the fast
// row coordinate and its two ends are confirmed with the exact test.
#include "mct_internal.h"

#include <algorithm>
#include <vector>

namespace {

using mct::kBufferDword3;

// Window of packed cells of one (tile, chunk, view) for the LDS-staged cone
// kernel: first cell (column c0, row r0) and extent w x h.  Columns lie in
// [0, n_u-2]; rows are not clipped to the detector (r0 may be negative, and
// r0 + h may pass n_v - 1), so every voxel of the tile and chunk, inside the
// mask or not, has its row coordinate inside the window.
struct BpWin {
    int16_t c0, r0, w, h;
};

struct BpArgs {
    int nx, nz, nu, nv, nviews;
    float pix, vh, cx, cy, cz, cos_rot, sin_rot;
    float sid, sdd, div_factor, det_h;
    float u0, du, inv_du, v0, dv, adv, vinv, vlim;
    float ua, ub, da, db;  // exact range-test intervals (see the file comment)
    int curved, cone, view_mode, apply_pm;
    const float *pm, *center_z;
    const float4 *vt;  // {cos, sin, dbeta, 0} of (angle - img_rot) per view
    const void *sino;
    const float *mask;  // IR only: ray mask multiplying the sinogram (or NULL)
    int layout;
    float *vol;
    int tiles_x, tiles_y, nzc, nxcd;
    // LDS-staged cone kernel: the plan's tables and the LDS window shape
    // (win == NULL: the direct kernel runs this job)
    const BpWin *win;
    const int2 *iv;
    int pitch, rows;
    // FBP view block: views [j0, j0 + nvb) of the scan, read from a packed
    // sinogram that holds only those views (j0 = 0, nvb = nviews: all).
    // accum 0: the volume is overwritten (masked voxels 0); 1: the block's
    // contributions are added to the volume in view order and masked voxels
    // are marked -10000; 2: as 1, then every voxel below -5000 is set to 0.
    int j0, nvb, accum;
    // LDS-staged kernel: first 16-slice chunk of the launch (its blocks cover
    // chunks zc0, zc0 + 1, ...)
    int zc0;
};

// Stored value of a voxel whose running sum is acc (alive: it passed the range
// tests of every view the launch decides on).
__device__ __forceinline__ float bp_result(const BpArgs &a, bool alive, float acc) {
    float v = alive ? acc : (a.accum ? -10000.f : 0.f);
    if (a.accum == 2 && v < -5000.f) v = 0.f;
    return v;
}

template <int DT> struct Taps;
template <> struct Taps<MCT_F32> {
    static constexpr int kQuadBytes = 16, kPairBytes = 8;
    using Raw = float4;
    __device__ __forceinline__ static Raw load_raw(const char *p) {
        return *reinterpret_cast<const float4 *>(p);
    }
    __device__ __forceinline__ static Raw load_buf(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    }
    __device__ __forceinline__ static float4 cvt(float4 q) { return q; }
    __device__ __forceinline__ static float4 quad(const void *base, size_t i) {
        return reinterpret_cast<const float4 *>(base)[i];
    }
    __device__ __forceinline__ static float4 quad_at(const char *p) {
        return *reinterpret_cast<const float4 *>(p);
    }
    __device__ __forceinline__ static float2 pair(const void *base, size_t i) {
        return reinterpret_cast<const float2 *>(base)[i];
    }
    __device__ __forceinline__ static float one(const void *base, size_t i) {
        return reinterpret_cast<const float *>(base)[i];
    }
    // ROWS: the taps {a, b-a} of one row element, and the QUAD cell of two
    // vertically neighbouring ones: {a, b-a, c-a, (d-c)-(b-a)}
    using Pair = float2;
    __device__ __forceinline__ static Pair load_pair(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ __forceinline__ static Raw quad_of(Pair p0, Pair p1) {
        return make_float4(p0.x, p0.y, p1.x - p0.x, p1.y - p0.y);
    }
};
template <> struct Taps<MCT_F16> {
    static constexpr int kQuadBytes = 8, kPairBytes = 4;
    using Raw = ushort4;
    __device__ __forceinline__ static Raw load_raw(const char *p) {
        return *reinterpret_cast<const ushort4 *>(p);
    }
    __device__ __forceinline__ static Raw load_buf(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_bit_cast(ushort4, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ __forceinline__ static float4 cvt(ushort4 h) {
        return make_float4(mct::f16_to_f32(h.x), mct::f16_to_f32(h.y), mct::f16_to_f32(h.z),
                           mct::f16_to_f32(h.w));
    }
    __device__ __forceinline__ static float4 quad(const void *base, size_t i) {
        return cvt(reinterpret_cast<const ushort4 *>(base)[i]);
    }
    __device__ __forceinline__ static float4 quad_at(const char *p) {
        return cvt(*reinterpret_cast<const ushort4 *>(p));
    }
    __device__ __forceinline__ static float2 pair(const void *base, size_t i) {
        ushort2 h = reinterpret_cast<const ushort2 *>(base)[i];
        return make_float2(mct::f16_to_f32(h.x), mct::f16_to_f32(h.y));
    }
    __device__ __forceinline__ static float one(const void *base, size_t i) {
        return mct::f16_to_f32(reinterpret_cast<const uint16_t *>(base)[i]);
    }
    using Pair = uint32_t;  // {a, b-a} as two FP16 halves, a in the low one
    __device__ __forceinline__ static Pair load_pair(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    }
    __device__ __forceinline__ static Raw quad_of(Pair p0, Pair p1) {
        const uint16_t a0 = (uint16_t)p0, d0 = (uint16_t)(p0 >> 16);
        const uint16_t a1 = (uint16_t)p1, d1 = (uint16_t)(p1 >> 16);
        return make_ushort4(a0, d0, mct::f32_to_f16_rn(mct::f16_to_f32(a1) - mct::f16_to_f32(a0)),
                            mct::f32_to_f16_rn(mct::f16_to_f32(d1) - mct::f16_to_f32(d0)));
    }
};
template <> struct Taps<MCT_BF16> {
    static constexpr int kQuadBytes = 8, kPairBytes = 4;
    using Raw = uint2;
    __device__ __forceinline__ static Raw load_raw(const char *p) {
        return *reinterpret_cast<const uint2 *>(p);
    }
    __device__ __forceinline__ static Raw load_buf(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ __forceinline__ static float4 cvt(uint2 w) {
        return make_float4(__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xffff0000u),
                           __uint_as_float(w.y << 16), __uint_as_float(w.y & 0xffff0000u));
    }
    __device__ __forceinline__ static float4 quad(const void *base, size_t i) {
        return cvt(reinterpret_cast<const uint2 *>(base)[i]);
    }
    __device__ __forceinline__ static float4 quad_at(const char *p) {
        return cvt(*reinterpret_cast<const uint2 *>(p));
    }
    __device__ __forceinline__ static float2 pair(const void *base, size_t i) {
        uint32_t w = reinterpret_cast<const uint32_t *>(base)[i];
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
    }
    __device__ __forceinline__ static float one(const void *base, size_t i) {
        return mct::bf16_to_f32(reinterpret_cast<const uint16_t *>(base)[i]);
    }
    using Pair = uint32_t;  // {a, b-a} as two BF16 halves, a in the low one
    __device__ __forceinline__ static Pair load_pair(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    }
    __device__ __forceinline__ static Raw quad_of(Pair p0, Pair p1) {
        typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
        typedef float f32x2 __attribute__((ext_vector_type(2)));
        const f32x2 ce = {__uint_as_float(p1 << 16) - __uint_as_float(p0 << 16),
                          __uint_as_float(p1 & 0xffff0000u) - __uint_as_float(p0 & 0xffff0000u)};
        // both differences rounded to nearest even in one v_cvt_pk_bf16_f32
        return make_uint2(p0, __builtin_bit_cast(uint32_t, __builtin_convertvector(ce, bf16x2)));
    }
};

// Bilinear value from packed coefficients {a, b-a, c-a, d-c-b+a}.
__device__ __forceinline__ float bilerp(float4 q, float ru, float rv) {
    return fmaf(rv, fmaf(ru, q.w, q.z), fmaf(ru, q.y, q.x));
}

// Logical tile of this workgroup: 16 x 16 columns (four 8 x 8 wavefront
// tiles) and one chunk of ZC slices.
__device__ __forceinline__ void decode_tile(const BpArgs &a, int b, int &ix, int &iy, int &zc_idx) {
    int t = b % (a.tiles_x * a.tiles_y);
    zc_idx = b / (a.tiles_x * a.tiles_y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    ix = (t % a.tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    iy = (t / a.tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
}

// In-plane position of voxel column (ix, iy) after ImageRotation (axial
// view), run_mgfbp.py:1089-1091, 1103-1104.
__device__ __forceinline__ void voxel_xy(const BpArgs &a, int ix, int iy, float &x, float &y) {
    float half = (float)(a.nx - 1) / 2.0f;
    float xa = a.pix * ((float)ix - half) + a.cx;
    float ya = (-a.pix) * ((float)iy - half) + a.cy;
    x = xa * a.cos_rot + ya * a.sin_rot;
    y = (-xa) * a.sin_rot + ya * a.cos_rot;
}

struct ColumnView {
    bool ok;
    uint32_t uf;
    float ru, mag, coef;
};

// Per (column, view) geometry of run_mgfbp.py:1122-1155 for a flat
// detector; the u range test is exact (ua <= u <= ub).
__device__ __forceinline__ ColumnView column_view(const BpArgs &a, float x, float y, float4 t) {
    ColumnView cv;
    float L = a.sid - x * t.x - y * t.y;
    cv.mag = a.sdd / L;
    float yr = (-x) * t.y + y * t.x;
    float u;
    float w;
    if (a.curved) {
        u = a.sdd * atan2f(yr, L);
        float q = x * t.y - y * t.x;
        w = __builtin_amdgcn_rcpf((L * L) + q * q);
    } else {
        u = cv.mag * yr;
        w = __builtin_amdgcn_rcpf(L * L);
    }
    cv.ok = u >= a.ua && u <= a.ub;
    float uidx = (u - a.u0) * a.inv_du;
    int uf = (int)uidx;  // uidx >= -tiny when ok: truncation is the floor
    uf = uf < 0 ? 0 : (uf > a.nu - 2 ? a.nu - 2 : uf);
    cv.uf = (uint32_t)uf;
    cv.ru = uidx - (float)uf;
    cv.coef = w * t.w;  // t.w = sid * dbeta * div_factor
    return cv;
}

// Exact v range test of slice iz: d = fl(fl(mag * z) - v0) in [da, db].
__device__ __forceinline__ bool v_exact(const BpArgs &a, float mag, int iz, float halfz) {
    float z = ((float)iz - halfz) * a.vh + a.cz;
    float d = mag * z - a.v0;
    return d >= a.da && d <= a.db;
}

// Slices [lo, hi] of the chunk starting at z0 (nk slices) whose projection
// passes the v test in this view.  t(k) = t0 + k dt is the fast row
// coordinate; the ends found from it are confirmed with the exact test.
__device__ __forceinline__ void v_interval(const BpArgs &a, float mag, int z0, int nk, float halfz,
                                           float t0, float dt, int &lo, int &hi) {
    const float vmax = (float)(a.nv - 1);
    const float rdt = __builtin_amdgcn_rcpf(dt);  // approximate: the ends are confirmed below
    float k_at0 = -t0 * rdt, k_atm = (vmax - t0) * rdt;
    float kf_lo = dt > 0.f ? k_at0 : k_atm, kf_hi = dt > 0.f ? k_atm : k_at0;
    // first k with t in range, last k with t in range (fast)
    kf_lo = fminf(fmaxf(kf_lo, -2.f), (float)nk + 2.f);
    kf_hi = fminf(fmaxf(kf_hi, -2.f), (float)nk + 2.f);
    int l = (int)ceilf(kf_lo);
    int h = (int)ceilf(kf_hi) - 1;
    // confirm the ends with the exact test (the fast coordinate is off by far
    // less than one slice, so each end moves by at most one)
    if (l - 1 >= 0 && l - 1 < nk && v_exact(a, mag, z0 + l - 1, halfz))
        l -= 1;
    else if (l >= 0 && l < nk && !v_exact(a, mag, z0 + l, halfz))
        l += 1;
    if (h + 1 >= 0 && h + 1 < nk && v_exact(a, mag, z0 + h + 1, halfz))
        h += 1;
    else if (h >= 0 && h < nk && !v_exact(a, mag, z0 + h, halfz))
        h -= 1;
    lo = l;
    hi = h;
}

// Cone beam, axial view, QUAD taps; ZC slices per thread, loads in batches
// of B.
template <int ZC, int B, int DT, bool ROWS>
__device__ __forceinline__ void cone_tile(const BpArgs &a, int block) {
    static_assert(ZC % B == 0, "batch must divide the slice chunk");
    int ix, iy, zci;
    decode_tile(a, block, ix, iy, zci);
    if (ix >= a.nx || iy >= a.nx) return;
    const int z0 = zci * ZC;
    const int nk = a.nz - z0 < ZC ? a.nz - z0 : ZC;
    float x, y;
    voxel_xy(a, ix, iy, x, y);
    const float halfz = (float)(a.nz - 1) / 2.0f;
    const float zfirst = ((float)z0 - halfz) * a.vh + a.cz;
    const float zstep_t = a.vh * a.vinv;
    const size_t plane = (size_t)a.nx * a.nx;
    float *out = a.vol + (size_t)iy * a.nx + ix;
    float acc[ZC];
#pragma unroll
    for (int k = 0; k < ZC; ++k) acc[k] = (a.accum && k < nk) ? out[(size_t)(z0 + k) * plane] : 0.f;
    int lo = 0, hi = nk - 1;
    // QUAD: cell rows -1 .. nv - 1; ROWS: element rows -1 .. nv (PAIR taps)
    const uint32_t rowbytes = (uint32_t)(a.nu - 1) * (ROWS ? Taps<DT>::kPairBytes : Taps<DT>::kQuadBytes);
    const size_t viewbytes = (size_t)(a.nv + (ROWS ? 2 : 1)) * rowbytes;
    const char *sino = reinterpret_cast<const char *>(a.sino);
    const float4 *vt = a.vt + a.j0;
    for (int j = 0; j < a.nvb; ++j) {
        const float4 t = vt[j];
        ColumnView cv = column_view(a, x, y, t);
        if (!cv.ok) {
            hi = -1;
            break;
        }
        const float t0 = (cv.mag * zfirst - a.v0) * a.vinv;
        const float dt = cv.mag * zstep_t;
        int l, h;
        v_interval(a, cv.mag, z0, nk, halfz, t0, dt, l, h);
        lo = l > lo ? l : lo;
        hi = h < hi ? h : hi;
        if (lo > hi) break;
        // one buffer resource per view (SGPRs), based at cell row 0: 32-bit
        // per-lane offsets
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(sino + (size_t)j * viewbytes + rowbytes), 0, (int)(viewbytes - rowbytes),
            kBufferDword3);
        // batches of B independent loads, issued before any of them is used
        if constexpr (ROWS) {  // the two rows of a cell: two PAIR loads
            const uint32_t uoff = cv.uf * (uint32_t)Taps<DT>::kPairBytes;
#pragma unroll
            for (int k0 = 0; k0 < ZC; k0 += B) {
                typename Taps<DT>::Pair p0[B], p1[B];
                float rv[B];
#pragma unroll
                for (int i = 0; i < B; ++i) {
                    float tk = fmaf((float)(k0 + i), dt, t0);
                    tk = __builtin_amdgcn_fmed3f(tk, 0.f, a.vlim);
                    uint32_t vf = (uint32_t)tk;
                    rv[i] = __builtin_amdgcn_fractf(tk);
                    p0[i] = Taps<DT>::load_pair(rsrc, __umul24(vf, rowbytes) + uoff);
                    p1[i] = Taps<DT>::load_pair(rsrc, __umul24(vf, rowbytes) + rowbytes + uoff);
                }
#pragma unroll
                for (int i = 0; i < B; ++i)
                    acc[k0 + i] = fmaf(cv.coef, bilerp(Taps<DT>::cvt(Taps<DT>::quad_of(p0[i], p1[i])), cv.ru, rv[i]),
                                       acc[k0 + i]);
            }
        } else {
            const uint32_t uoff = cv.uf * (uint32_t)Taps<DT>::kQuadBytes;
#pragma unroll
            for (int k0 = 0; k0 < ZC; k0 += B) {
                typename Taps<DT>::Raw raw[B];
                float rv[B];
#pragma unroll
                for (int i = 0; i < B; ++i) {
                    float tk = fmaf((float)(k0 + i), dt, t0);
                    tk = __builtin_amdgcn_fmed3f(tk, 0.f, a.vlim);
                    uint32_t vf = (uint32_t)tk;
                    rv[i] = __builtin_amdgcn_fractf(tk);
                    raw[i] = Taps<DT>::load_buf(rsrc, __umul24(vf, rowbytes) + uoff);
                }
#pragma unroll
                for (int i = 0; i < B; ++i)
                    acc[k0 + i] = fmaf(cv.coef, bilerp(Taps<DT>::cvt(raw[i]), cv.ru, rv[i]), acc[k0 + i]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < ZC; ++k)
        if (k < nk) out[(size_t)(z0 + k) * plane] = bp_result(a, k >= lo && k <= hi, acc[k]);
}

template <int ZC, int B, int DT, bool ROWS>
__global__ __launch_bounds__(256) void bp_cone_axial_kernel(BpArgs a) {
    cone_tile<ZC, B, DT, ROWS>(a, mct::xcd_remap((int)blockIdx.x, a.tiles_x * a.tiles_y * a.nzc, a.nxcd));
}

// ---------------------------------------------------------------------------
// LDS-staged cone back-projector (axial view, flat or curved detector, ROWS
// or QUAD taps; the LDS window holds QUAD cells either way).  A workgroup owns a 16 x 16-column tile and a chunk of 16 slices
// (one column and all 16 slices per thread).  In one view the tile's voxels
// read packed cells from a small window of the detector; the window of every
// (tile, chunk, view) is tabulated once per plan from the tile's corners: a
// voxel's detector coordinates are linear-fractional in its position (and the
// row coordinate linear in z), so their extremes over the tile are taken at
// the corners; a margin of 1/64 of a cell (plus 2^-18 of the coordinate's
// magnitude) on each side covers the kernel's FP32 rounding, which stays
// below 1e-3 of a cell at the sizes the plan accepts.  The
// workgroup copies the window of view j+1 into one LDS buffer while it
// interpolates view j from the other, so a packed cell crosses the vector
// memory path once per tile and view instead of once per voxel.  The mask (a
// voxel that leaves the detector in any view is 0) comes from the plan's
// per-column slice intervals, decided with the same exact tests as the
// direct kernel.

constexpr int kLdsTile = 16;  // columns per side of a workgroup tile (decode_tile)
constexpr int kLdsZC = 16;    // slices per thread
constexpr int kViewBatch = 32;  // per-view table entries staged in LDS at a time (a power of 2)
static_assert((kViewBatch & (kViewBatch - 1)) == 0 && kViewBatch <= 64, "view batch");

struct FastView {
    uint32_t uf;
    float ru, mag, coef;
};

// Per (column, view) interpolation geometry of run_mgfbp.py:1122-1155 for the
// staged kernel.  The mask is decided by the plan's exact slice intervals,
// so 1/L comes from the hardware reciprocal (1 ulp) instead of a division.
__device__ __forceinline__ FastView column_view_fast(const BpArgs &a, float x, float y, float4 t) {
    FastView cv;
    const float L = a.sid - x * t.x - y * t.y;
    const float rL = __builtin_amdgcn_rcpf(L);
    cv.mag = a.sdd * rL;
    const float yr = (-x) * t.y + y * t.x;
    float u, w;
    if (a.curved) {
        u = a.sdd * atan2f(yr, L);
        const float q = x * t.y - y * t.x;
        w = __builtin_amdgcn_rcpf((L * L) + q * q);
    } else {
        u = cv.mag * yr;
        w = rL * rL;
    }
    const float uidx = (u - a.u0) * a.inv_du;
    // A column the staged kernel interpolates passes every view's u test, so
    // 0 <= uidx <= nu - 2: truncation is the floor and fract(uidx) equals
    // uidx - floor(uidx).  Other columns only need an address, which the
    // window lookup clamps.
    cv.uf = (uint32_t)(int)uidx;
    cv.ru = __builtin_amdgcn_fractf(uidx);
    cv.coef = w * t.w;  // t.w = sid * dbeta * div_factor
    return cv;
}

// d = fl(fl(mag * z) - v[0]) of slice iz, the quantity of the exact v test.
__device__ __forceinline__ float v_coord(const BpArgs &a, float mag, int iz, float halfz) {
    float z = ((float)iz - halfz) * a.vh + a.cz;
    return mag * z - a.v0;
}

// Per image column: the slices that pass the range tests of every view,
// [lo, hi] (lo > hi: none).  d(iz) is non-decreasing in iz (mag > 0, vh > 0),
// so each view's passing slices form an interval whose ends are found by
// bisection with the exact test.
__global__ __launch_bounds__(256) void bp_column_iv_kernel(BpArgs a, int2 *__restrict__ iv,
                                                           int *__restrict__ kept) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y;
    if (ix >= a.nx) return;
    float x, y;
    voxel_xy(a, ix, iy, x, y);
    const float halfz = (float)(a.nz - 1) / 2.0f;
    int lo = 0, hi = a.nz - 1;
    for (int j = 0; j < a.nviews && lo <= hi; ++j) {
        const ColumnView cv = column_view(a, x, y, a.vt[j]);
        if (!cv.ok) {
            lo = 1;
            hi = 0;
            break;
        }
        int l = lo, r = hi + 1;  // first slice with d >= da
        while (l < r) {
            int m = (l + r) >> 1;
            if (v_coord(a, cv.mag, m, halfz) >= a.da)
                r = m;
            else
                l = m + 1;
        }
        const int nlo = l;
        l = lo;
        r = hi + 1;  // first slice with d > db
        while (l < r) {
            int m = (l + r) >> 1;
            if (v_coord(a, cv.mag, m, halfz) > a.db)
                r = m;
            else
                l = m + 1;
        }
        lo = nlo;
        hi = l - 1;
    }
    iv[(size_t)iy * a.nx + ix] = make_int2(lo, hi);
    // kept[0] = nz - (lowest kept slice), kept[1] = highest kept slice + 1
    // over all columns (0: none), reduced over the wavefront first
    int k0 = lo <= hi ? a.nz - lo : 0, k1 = lo <= hi ? hi + 1 : 0;
    for (int off = 32; off > 0; off >>= 1) {
        k0 = max(k0, __shfl_xor(k0, off));
        k1 = max(k1, __shfl_xor(k1, off));
    }
    if ((threadIdx.x & 63) == 0 && k1 > 0) {
        atomicMax(&kept[0], k0);
        atomicMax(&kept[1], k1);
    }
}

// Window table, FP64 corner evaluation; maxes[0..1] collect the largest w, h.
// The rows span every slice position of the chunk, kLdsZC of them even in a
// last, partial chunk (the kernel computes those slices and discards them),
// and are not clipped to the detector.  A tile with a corner at or behind
// the source (L <= 0) gets an oversized window, which makes the launch fall
// back to the direct kernel.
__global__ __launch_bounds__(64) void bp_window_kernel(BpArgs a, BpWin *__restrict__ win,
                                                       int *__restrict__ maxes) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nviews) return;
    const int tile = blockIdx.y, zci = blockIdx.z;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int ixs[2] = {tx * kLdsTile, min(tx * kLdsTile + kLdsTile - 1, a.nx - 1)};
    const int iys[2] = {ty * kLdsTile, min(ty * kLdsTile + kLdsTile - 1, a.nx - 1)};
    const int iz0 = zci * kLdsZC, iz1 = iz0 + kLdsZC - 1;
    const double half = (a.nx - 1) / 2.0, halfz = (a.nz - 1) / 2.0;
    const double c = a.vt[j].x, s = a.vt[j].y;
    const double zs[2] = {(iz0 - halfz) * a.vh + a.cz, (iz1 - halfz) * a.vh + a.cz};
    double umin = 1e300, umax = -1e300, tmin = 1e300, tmax = -1e300;
    bool bad = false;
    for (int p = 0; p < 2; ++p) {
        for (int q = 0; q < 2; ++q) {
            const double xa = (double)a.pix * (ixs[p] - half) + a.cx;
            const double ya = -(double)a.pix * (iys[q] - half) + a.cy;
            const double x = xa * a.cos_rot + ya * a.sin_rot;
            const double y = -xa * a.sin_rot + ya * a.cos_rot;
            const double L = a.sid - x * c - y * s;
            if (!(L > 0.0)) bad = true;
            const double yr = -x * s + y * c;
            const double mag = a.sdd / L;
            const double u = a.curved ? a.sdd * atan2(yr, L) : mag * yr;
            const double ui = (u - a.u0) * a.inv_du;
            umin = fmin(umin, ui);
            umax = fmax(umax, ui);
            for (int r = 0; r < 2; ++r) {
                const double t = (mag * zs[r] - a.v0) * a.vinv;
                tmin = fmin(tmin, t);
                tmax = fmax(tmax, t);
            }
        }
    }
    const double cmax = a.nu - 2;
    // cell c serves coordinates [c, c + 1): the window holds the cells of
    // [min - e, max + e], e covering the kernel's FP32 rounding (1/64 of a
    // cell, plus 2^-18 of the magnitude for very large detectors)
    const double eu = 1.0 / 64.0 + fmax(fabs(umin), fabs(umax)) * 0x1p-18;
    const double et = 1.0 / 64.0 + fmax(fabs(tmin), fabs(tmax)) * 0x1p-18;
    const double cl = fmin(fmax(floor(umin - eu), 0.0), cmax);
    const double ch = fmin(fmax(floor(umax + eu), 0.0), cmax);
    const double rl = floor(tmin - et);
    const double rh = floor(tmax + et);
    const bool fits = !bad && umin <= umax && tmin <= tmax && rl >= -32768.0 && rh <= 32767.0;
    int w = 32767, h = 32767;
    if (fits) {
        w = ch >= cl ? (int)(ch - cl) + 1 : 0;
        h = (int)fmin(rh - rl + 1.0, 32767.0);
    }
    BpWin e;
    e.c0 = (int16_t)cl;
    e.r0 = fits ? (int16_t)rl : (int16_t)0;
    e.w = (int16_t)w;
    e.h = (int16_t)h;
    win[((size_t)tile * a.nzc + zci) * a.nviews + j] = e;
    atomicMax(&maxes[0], w);
    atomicMax(&maxes[1], h);
}

// One workgroup tile (tile, chunk) = block of job a.  a.pitch x a.rows: the
// LDS window shape (largest w and h of the job's table); NQ: staging slots
// per thread, NQ * 256 >= pitch * rows; smem: two buffers of pitch * rows
// packed cells; ACC: the launch may be a view block (a.accum != 0), else the
// volume is overwritten.  Every thread of the workgroup must call this
// (barriers).
template <int DT, int NQ, bool ACC, bool ROWS>
__device__ __forceinline__ void cone_lds_tile(const BpArgs &a, int block, unsigned char *smem) {
    using T = Taps<DT>;
    using Raw = typename T::Raw;
    Raw *lds = reinterpret_cast<Raw *>(smem);
    const int pitch = a.pitch, rows = a.rows;
    const BpWin *__restrict__ win = a.win;
    const int2 *__restrict__ iv = a.iv;
    const int cells = pitch * rows;
    // Tiles of a chunk in Z order when the tile grid is a square power of
    // two, so that the workgroups resident at one time form compact blocks
    // (their detector footprints overlap more than those of a row strip).
    const int ntile = a.tiles_x * a.tiles_y;
    const int tz = block % ntile, zci = a.zc0 + block / ntile;
    if (zci >= a.nzc) return;  // (uniform over the workgroup: before any barrier)
    int tx, ty;
    if (a.tiles_x == a.tiles_y && (a.tiles_x & (a.tiles_x - 1)) == 0) {
        auto even_bits = [](uint32_t v) {
            v &= 0x55555555u;
            v = (v | (v >> 1)) & 0x33333333u;
            v = (v | (v >> 2)) & 0x0f0f0f0fu;
            v = (v | (v >> 4)) & 0x00ff00ffu;
            v = (v | (v >> 8)) & 0x0000ffffu;
            return (int)v;
        };
        tx = even_bits((uint32_t)tz);
        ty = even_bits((uint32_t)tz >> 1);
    } else {
        tx = tz % a.tiles_x;
        ty = tz / a.tiles_x;
    }
    const int tile = ty * a.tiles_x + tx;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ix = tx * 16 + (wave & 1) * 8 + (lane & 7);
    const int iy = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool inside = ix < a.nx && iy < a.nx;
    const int z0 = zci * kLdsZC;
    const int nk = a.nz - z0 < kLdsZC ? a.nz - z0 : kLdsZC;
    int lo = 0, hi = -1;
    if (inside) {
        const int2 v = iv[(size_t)iy * a.nx + ix];
        lo = v.x - z0 > 0 ? v.x - z0 : 0;
        hi = v.y - z0 < nk - 1 ? v.y - z0 : nk - 1;
    }
    const bool alive = lo <= hi;
    float acc[kLdsZC];
#pragma unroll
    for (int k = 0; k < kLdsZC; ++k) acc[k] = 0.f;
    if (ACC && a.accum && inside) {  // a view block adds to the volume
        const size_t plane = (size_t)a.nx * a.nx;
        const float *in = a.vol + (size_t)iy * a.nx + ix;
#pragma unroll
        for (int k = 0; k < kLdsZC; ++k)
            if (k < nk) acc[k] = in[(size_t)(z0 + k) * plane];
    }

    if (__syncthreads_or(alive)) {
        const int tid = threadIdx.x;
        // QUAD: cell rows -1 .. nv - 1; ROWS: element rows -1 .. nv
        const uint32_t rowbytes = (uint32_t)(a.nu - 1) * (ROWS ? T::kPairBytes : T::kQuadBytes);
        const size_t viewbytes = (size_t)(a.nv + (ROWS ? 2 : 1)) * rowbytes;
        const char *sino = reinterpret_cast<const char *>(a.sino);
        // (a whole-scan launch reads the fields as given: j0 = 0, nvb = nviews)
        const BpWin *wt = win + ((size_t)tile * a.nzc + zci) * a.nviews + (ACC ? a.j0 : 0);
        const float4 *vt = ACC ? a.vt + a.j0 : a.vt;
        const int nvb = ACC ? a.nvb : a.nviews;
        // A window's rows [r0, r0 + h) are not clipped to the detector, so no
        // row coordinate of the tile leaves it and the slice loop needs no
        // clamp.  Of those rows, the packed cell rows [-1, nv - 1] (the
        // detector's cells and the two zero-padded ones around them, where a
        // voxel inside the mask can land by the rounding of its row
        // coordinate) are loaded; other rows keep whatever the buffer held
        // and are read only for voxels outside the mask, whose sums are
        // discarded.
        auto first_row = [&](const BpWin &w) { return w.r0 > -1 ? (int)w.r0 : -1; };
        auto row_count = [&](const BpWin &w) {
            const int end = (int)w.r0 + (int)w.h < a.nv ? (int)w.r0 + (int)w.h : a.nv;
            return end > first_row(w) ? end - first_row(w) : 0;
        };
        // QUAD: staging slot q of this thread is entry e = tid + 256 q of the
        // rows loaded, i.e. (row e / pitch, column e % pitch) from the first
        // loaded cell; its byte offset from that cell is the same in every
        // view (a slot is in the first n rows exactly when soff < n *
        // rowbytes; slots past the buffer get an offset no row count
        // reaches).
        // ROWS: strip q of this thread (NQ / 2 of them) is s = tid + 256 q:
        // window column s % pitch, cell rows 4 (s / pitch) .. + 3; it loads
        // the five element rows of its four cells (8 B each for FP32: 10 B
        // per cell against QUAD's 16) and forms the cells in registers.
        constexpr int NS = ROWS ? NQ / 2 : 1;
        constexpr int NSLOT = ROWS ? 1 : NQ;
        uint32_t soff[NSLOT];
        uint32_t goff[NS];
        int lidx[NS], srow0[NS];
        if constexpr (ROWS) {
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const int e = tid + 256 * q;
                const int g = e / pitch, c = e - g * pitch;
                srow0[q] = 4 * g < rows ? 4 * g : 0x7fffffff;
                goff[q] = (uint32_t)(4 * g) * rowbytes + (uint32_t)c * T::kPairBytes;
                lidx[q] = 4 * g * pitch + c;
            }
        } else {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int e = tid + 256 * q;
                const int r = e / pitch, c = e - r * pitch;
                soff[q] = e < cells ? (uint32_t)r * rowbytes + (uint32_t)c * T::kQuadBytes : 0xffffffffu;
            }
        }
        Raw stage[NSLOT];
        typename T::Pair pst[NS][5];
        auto fetch = [&](int j, const BpWin &w) {
            const int nr = row_count(w);
            // first loaded row: stored row first_row + 1 of the view
            const size_t origin = (size_t)(first_row(w) + 1) * rowbytes +
                                  (size_t)w.c0 * (ROWS ? T::kPairBytes : T::kQuadBytes);
            if constexpr (ROWS) {
                // the resource ends after element row nr (the last the loaded
                // cells read): loads past it return zeros
                const size_t rest = viewbytes - origin, need = (size_t)(nr + 1) * rowbytes;
                const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(sino + (size_t)j * viewbytes + origin), 0, (int)(need < rest ? need : rest),
                    kBufferDword3);
#pragma unroll
                for (int q = 0; q < NS; ++q)
                    if (srow0[q] < nr) {
#pragma unroll
                        for (int k = 0; k < 5; ++k) pst[q][k] = T::load_pair(rsrc, goff[q] + (uint32_t)k * rowbytes);
                    }
            } else {
                const uint32_t lim = (uint32_t)nr * rowbytes;
                const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(sino + (size_t)j * viewbytes + origin), 0, (int)(viewbytes - origin),
                    kBufferDword3);
                // (slots past the rows to load skip their load: a range-checked
                // load that returns zeros still occupies the address unit, and
                // made the launch 3% slower)
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    if (soff[q] < lim) stage[q] = T::load_buf(rsrc, soff[q]);
            }
        };
        auto put = [&](int b, const BpWin &w) {
            const int nr = row_count(w);
            Raw *buf = lds + b * cells + (first_row(w) - (int)w.r0) * pitch;
            if constexpr (ROWS) {
#pragma unroll
                for (int q = 0; q < NS; ++q)
                    if (srow0[q] < nr) {
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (srow0[q] + k < nr) buf[lidx[q] + k * pitch] = T::quad_of(pst[q][k], pst[q][k + 1]);
                    }
            } else {
                const uint32_t lim = (uint32_t)nr * rowbytes;
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    if (soff[q] < lim) buf[tid + 256 * q] = stage[q];
            }
        };
        // Per-view table entries pass through LDS in batches of kViewBatch
        // views: s_vt[i] = {cos, sin, dbeta, weight} of view j0 + i, s_win[i]
        // = window of view j0 + i + 1.  (Read from global memory inside the
        // loop they would be vector loads queued behind the window prefetch,
        // and waiting for them would wait for the prefetch.)  Batches of 32
        // rather than 64 keep the static LDS at 1 KB, which with 12.8 KB
        // windows lets 6 workgroups share a CU instead of 5.
        __shared__ float4 s_vt[kViewBatch];
        __shared__ BpWin s_win[kViewBatch];
        auto load_tables = [&](int j0) {
            if (tid < kViewBatch) {
                if (j0 + tid < nvb) s_vt[tid] = vt[j0 + tid];
            } else if (tid >= 64 && tid < 64 + kViewBatch) {
                const int jj = j0 + 1 + (tid - 64);
                if (jj < nvb) s_win[tid - 64] = wt[jj];
            }
        };
        float x, y;
        voxel_xy(a, ix, iy, x, y);
        const float halfz = (float)(a.nz - 1) / 2.0f;
        const float zfirst = ((float)z0 - halfz) * a.vh + a.cz;
        const float zstep_t = a.vh * a.vinv;
        const uint32_t row_stride = (uint32_t)pitch * (uint32_t)sizeof(Raw);  // bytes per window row

        load_tables(0);
        BpWin wc = wt[0];
        fetch(0, wc);
        put(0, wc);
        __syncthreads();
        for (int j = 0; j < nvb; ++j) {
            if (j && (j & (kViewBatch - 1)) == 0) {  // every wave has finished the previous batch's entries
                load_tables(j);
                __syncthreads();
            }
            const bool more = j + 1 < nvb;
            BpWin wn = wc;
            if (more) {
                wn = s_win[j & (kViewBatch - 1)];
                fetch(j + 1, wn);
            }
            if (alive) {
                const FastView cv = column_view_fast(a, x, y, s_vt[j & (kViewBatch - 1)]);
                // row coordinate relative to the window's first row: in
                // [0, h) for every slice of the chunk (the window spans the
                // tile's rows with a rounding margin on either side)
                const float t0 = (cv.mag * zfirst - a.v0) * a.vinv - (float)(int)wc.r0;
                const float dt = cv.mag * zstep_t;
                // (an unsigned minimum clamps both ends: a negative index wraps high)
                const uint32_t lc = min((uint32_t)((int)cv.uf - (int)wc.c0), (uint32_t)((int)wc.w - 1));
                const unsigned char *src =
                    smem + (uint32_t)((j & 1) * cells + lc) * (uint32_t)sizeof(Raw);
                // reads in batches of 4, all in flight before the first is used
                // (the scheduling barrier keeps the compiler from pairing
                // each read with its use: 22.3-22.4 ms against 22.7 ms at
                // 512^2 x 256 / 1080 views; batches of 8 need 106 VGPRs, one
                // wavefront per SIMD fewer, and measured 23.5 ms)
#pragma unroll
                for (int k0 = 0; k0 < kLdsZC; k0 += 4) {
                    Raw q[4];
                    float rv[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float tk = fmaf((float)(k0 + i), dt, t0);
                        const uint32_t vf = (uint32_t)tk;
                        rv[i] = __builtin_amdgcn_fractf(tk);
                        q[i] = *reinterpret_cast<const Raw *>(src + __umul24(vf, row_stride));
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        acc[k0 + i] = fmaf(cv.coef, bilerp(T::cvt(q[i]), cv.ru, rv[i]), acc[k0 + i]);
                }
            }
            if (more) put((j + 1) & 1, wn);
            __syncthreads();
            wc = wn;
        }
    }
    if (inside) {
        const size_t plane = (size_t)a.nx * a.nx;
        float *out = a.vol + (size_t)iy * a.nx + ix;
#pragma unroll
        for (int k = 0; k < kLdsZC; ++k) {
            const bool ok = k >= lo && k <= hi;
            if (k < nk) out[(size_t)(z0 + k) * plane] = ACC ? bp_result(a, ok, acc[k]) : (ok ? acc[k] : 0.f);
        }
    }
}

// Zeroes n floats (the slices a whole-scan launch leaves out).
__global__ __launch_bounds__(256) void bp_fill_zero_kernel(float *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// NQ = 2 (the windows of common geometries): at most 80 VGPRs, so 6
// wavefronts fit a SIMD when the LDS allows 6 workgroups per CU (at
// 512^2 x 256 / 1080 views: 18.1 ms against 18.7-18.8 ms with 5); wider
// windows keep the compiler's register choice (80 would spill).
template <int DT, int NQ, bool ACC, bool ROWS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NQ <= 2 ? 6 : 1)))
void bp_cone_lds_kernel(BpArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    // (the grid covers the launch's chunks only: tiles x chunks from zc0)
    cone_lds_tile<DT, NQ, ACC, ROWS>(a, mct::xcd_remap((int)blockIdx.x, (int)gridDim.x, a.nxcd), smem);
}

// Fan beam, axial view, PAIR taps; slice z reads detector row z.
template <int ZC, int DT>
__device__ __forceinline__ void fan_tile(const BpArgs &a, int block) {
    int ix, iy, zci;
    decode_tile(a, block, ix, iy, zci);
    if (ix >= a.nx || iy >= a.nx) return;
    const int z0 = zci * ZC;
    const int nk = a.nz - z0 < ZC ? a.nz - z0 : ZC;
    float x, y;
    voxel_xy(a, ix, iy, x, y);
    const size_t plane = (size_t)a.nx * a.nx;
    float *out = a.vol + (size_t)iy * a.nx + ix;
    float acc[ZC];
#pragma unroll
    for (int k = 0; k < ZC; ++k) acc[k] = (a.accum && k < nk) ? out[(size_t)(z0 + k) * plane] : 0.f;
    bool alive = true;
    const int ncol = a.nu - 1;
    const size_t row_stride = (size_t)a.nvb * ncol;
    const float4 *vt = a.vt + a.j0;
    for (int j = 0; j < a.nvb; ++j) {
        const float4 t = vt[j];
        ColumnView cv = column_view(a, x, y, t);
        if (!cv.ok) {
            alive = false;
            break;
        }
        const size_t base = (size_t)z0 * row_stride + (size_t)j * ncol + cv.uf;
#pragma unroll
        for (int k = 0; k < ZC; ++k) {
            if (k < nk) {
                float2 p = Taps<DT>::pair(a.sino, base + (size_t)k * row_stride);
                acc[k] = fmaf(cv.coef, fmaf(cv.ru, p.y, p.x), acc[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < ZC; ++k)
        if (k < nk) out[(size_t)(z0 + k) * plane] = bp_result(a, alive, acc[k]);
}

template <int ZC, int DT>
__global__ __launch_bounds__(256) void bp_fan_axial_kernel(BpArgs a) {
    fan_tile<ZC, DT>(a, mct::xcd_remap((int)blockIdx.x, a.tiles_x * a.tiles_y * a.nzc, a.nxcd));
}

// Back-projection of many volumes / slice stacks of different geometry in one
// launch.  The work items (one 16 x 16-column x ZC-slice tile of one job)
// are numbered job after job; offsets[j] is the first item of job j
// (offsets[njobs] = total).  A fixed grid walks the items in a static round
// robin, so the launch is deterministic and needs no counters (it can be
// captured in a graph and replayed).  Cone jobs with window tables run the
// LDS-staged tile (NQ staging slots; NQ = 0: none in this batch), the others
// the direct ones.  An item's job is uniform over the workgroup, so the
// staged tile's barriers are reached by every thread.
template <int DT, int NQ>
__global__ __launch_bounds__(256) void bp_batch_kernel(const BpArgs *__restrict__ jobs,
                                                       const int *__restrict__ offsets, int njobs,
                                                       int total) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = njobs - 1;  // last job with offsets[job] <= item
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (offsets[mid] <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BpArgs a = jobs[lo];
        const int local = item - offsets[lo];
        if constexpr (NQ > 0) {
            if (a.cone && a.win) {
                if (a.layout == MCT_ROWS)
                    cone_lds_tile<DT, NQ, false, true>(a, local, smem);
                else
                    cone_lds_tile<DT, NQ, false, false>(a, local, smem);
                continue;
            }
        }
        if (a.cone && a.layout == MCT_ROWS)
            cone_tile<16, 16, DT, true>(a, local);
        else if (a.cone)
            cone_tile<16, 16, DT, false>(a, local);
        else
            fan_tile<8, DT>(a, local);
    }
}

// Voxel position for any ReconViewMode, run_mgfbp.py:1086-1104.
__device__ __forceinline__ void voxel_xyz(const BpArgs &a, int ix, int iy, int iz, float czv,
                                          float &x, float &y, float &z) {
    float half = (float)(a.nx - 1) / 2.0f;
    float halfz = (float)(a.nz - 1) / 2.0f;
    float xa, ya;
    if (a.view_mode == 1) {
        xa = a.pix * ((float)ix - half) + a.cx;
        ya = (-a.pix) * ((float)iy - half) + a.cy;
        z = ((float)iz - halfz) * a.vh + czv;
    } else if (a.view_mode == 2) {
        xa = a.pix * ((float)ix - half) + a.cx;
        z = (-a.pix) * ((float)iy - half) + czv;
        ya = (-((float)iz - halfz)) * a.vh + a.cy;
    } else {
        z = (-a.pix) * ((float)iy - half) + czv;
        ya = (-a.pix) * ((float)ix - half) + a.cy;
        xa = ((float)iz - halfz) * a.vh + a.cx;
    }
    x = xa * a.cos_rot + ya * a.sin_rot;
    y = (-xa) * a.sin_rot + ya * a.cos_rot;
}

// Generic FBP back-projection: one thread per voxel; every view mode and the
// P-matrix branch (run_mgfbp.py:1133-1141, 1165-1168), QUAD/PAIR/PLAIN taps.
template <int DT>
__global__ __launch_bounds__(256) void bp_generic_kernel(BpArgs a) {
    int64_t n = (int64_t)a.nz * a.nx * a.nx;
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    int ix = (int)(id % a.nx);
    int iy = (int)((id / a.nx) % a.nx);
    int iz = (int)(id / ((int64_t)a.nx * a.nx));
    float x, y, z;
    voxel_xyz(a, ix, iy, iz, a.cz, x, y, z);
    const int ncol = a.nu - 1;
    float acc = a.accum ? a.vol[id] : 0.f;
    bool alive = true;
    // jl: view within the packed block; j = a.j0 + jl: view of the scan
    for (int jl = 0; jl < a.nvb; ++jl) {
        const int j = a.j0 + jl;
        const float4 t = a.vt[j];
        float L = a.sid - x * t.x - y * t.y;
        float mag, uidx;
        if (!a.apply_pm) {
            mag = a.sdd / L;
            float yr = (-x) * t.y + y * t.x;
            float u = a.curved ? a.sdd * atan2f(yr, L) : mag * yr;
            uidx = (u - a.u0) / a.du;
        } else {
            const float *P = a.pm + 12 * j;
            mag = 1.0f / (P[8] * x + P[9] * y + P[10] * z + P[11] * 1.0f);
            uidx = (P[0] * x + P[1] * y + P[2] * z + P[3] * 1.0f) * mag;
        }
        if (uidx < 0.f || uidx + 1.f > (float)(a.nu - 1)) {
            alive = false;
            break;
        }
        int uf = (int)floorf(uidx);
        uf = uf > a.nu - 2 ? a.nu - 2 : uf;
        float ru = uidx - (float)uf;
        float w;
        if (a.curved) {
            float q = x * t.y - y * t.x;
            w = 1.0f / ((L * L) + q * q);
        } else {
            w = 1.0f / (L * L);
        }
        float coef = (a.sid * w) * t.z * a.div_factor;
        if (a.cone) {
            float vidx;
            if (!a.apply_pm) {
                vidx = (mag * z - a.v0) / a.det_h * a.adv / a.dv;
            } else {
                const float *P = a.pm + 12 * j;
                vidx = (P[4] * x + P[5] * y + P[6] * z + P[7] * 1.0f) * mag;
            }
            float vff = floorf(vidx);
            if (vff < 0.f || vff + 1.f > (float)(a.nv - 1)) {
                alive = false;
                break;
            }
            int vf = (int)vff;
            float rv = vidx - vff;
            if (a.layout == MCT_QUAD) {
                float4 q = Taps<DT>::quad(a.sino, ((size_t)jl * (a.nv + 1) + vf + 1) * ncol + uf);
                acc += coef * bilerp(q, ru, rv);
            } else if (a.layout == MCT_ROWS) {
                const size_t r0 = ((size_t)jl * (a.nv + 2) + vf + 1) * ncol + uf;
                const typename Taps<DT>::Pair *pp = reinterpret_cast<const typename Taps<DT>::Pair *>(a.sino);
                float4 q = Taps<DT>::cvt(Taps<DT>::quad_of(pp[r0], pp[r0 + ncol]));
                acc += coef * bilerp(q, ru, rv);
            } else {
                size_t r0 = ((size_t)vf * a.nvb + jl) * a.nu + uf;
                size_t r1 = r0 + (size_t)a.nvb * a.nu;
                float p0 = Taps<DT>::one(a.sino, r0) * (1.f - ru) + Taps<DT>::one(a.sino, r0 + 1) * ru;
                float p1 = Taps<DT>::one(a.sino, r1) * (1.f - ru) + Taps<DT>::one(a.sino, r1 + 1) * ru;
                acc += coef * ((1.f - rv) * p0 + rv * p1);
            }
        } else {
            float s0, s1;
            if (a.layout == MCT_PAIR) {
                float2 p = Taps<DT>::pair(a.sino, ((size_t)iz * a.nvb + jl) * ncol + uf);
                s0 = p.x;
                s1 = p.x + p.y;
            } else {
                size_t r0 = ((size_t)iz * a.nvb + jl) * a.nu + uf;
                s0 = Taps<DT>::one(a.sino, r0);
                s1 = Taps<DT>::one(a.sino, r0 + 1);
            }
            acc += coef * ((1.f - ru) * s0 + ru * s1);
        }
    }
    a.vol[id] = bp_result(a, alive, acc);
}

// IR back-projection over all views (run_mgfbp_iterative_recon_IRN_CGLS.py:
// 372-481): unfiltered, weight = ray length per voxel * pix; a voxel whose u
// leaves the detector in any view is 0 (truncation flag); v off the detector
// adds 0.  PLAIN FP32 sinogram [rows][views][nu]; optional mask multiplies
// each tap.  ZC slices per thread.
// FLATZ: fan beam with voxel height 0, where every slice sits at z = the
// view's centre, so the ray-length weight is one value per (column, view).
template <int ZC, bool FLATZ>
__global__ __launch_bounds__(256) void bp_ir_axial_kernel(BpArgs a) {
    int ix, iy, zci;
    decode_tile(a, mct::xcd_remap((int)blockIdx.x, a.tiles_x * a.tiles_y * a.nzc, a.nxcd), ix, iy, zci);
    if (ix >= a.nx || iy >= a.nx) return;
    const int z0 = zci * ZC;
    const int nk = a.nz - z0 < ZC ? a.nz - z0 : ZC;
    float x, y;
    voxel_xy(a, ix, iy, x, y);
    const float halfz = (float)(a.nz - 1) / 2.0f;
    float zb[ZC], acc[ZC];
#pragma unroll
    for (int k = 0; k < ZC; ++k) {
        zb[k] = ((float)(z0 + k) - halfz) * a.vh;
        acc[k] = 0.f;
    }
    bool alive = true;
    const float *sg = reinterpret_cast<const float *>(a.sino);
    const size_t row_stride = (size_t)a.nviews * a.nu;
    const float vmax = (float)(a.nv - 1);
    for (int j = 0; j < a.nviews; ++j) {
        const float4 t = a.vt[j];
        const float czv = a.center_z ? a.center_z[j] : a.cz;
        float L = a.sid - x * t.x - y * t.y;
        float dxs = x - a.sid * t.x;
        float dys = y - a.sid * t.y;
        bool dflag = fabsf(dxs) > fabsf(dys);
        float den = dflag ? fabsf(dxs) : fabsf(dys);
        float mag = a.sdd / L;
        float yr = (-x) * t.y + y * t.x;
        float u = a.curved ? a.sdd * atan2f(yr, L) : mag * yr;
        float uidx = (u - a.u0) / a.du;
        if (uidx < 0.f || uidx + 1.f > (float)(a.nu - 1)) {
            alive = false;
            break;
        }
        int uf = (int)floorf(uidx);
        uf = uf > a.nu - 2 ? a.nu - 2 : uf;
        float ru = uidx - (float)uf;
        const size_t vb = (size_t)j * a.nu + uf;
        // (zb[k] = +-0 when FLATZ: z = czv exactly in every slice)
        const float wgt_view = FLATZ ? sqrtf(dxs * dxs + dys * dys + czv * czv) / den : 0.f;
#pragma unroll
        for (int k = 0; k < ZC; ++k) {
            if (k >= nk) continue;
            float z = zb[k] + czv;
            float wgt = FLATZ ? wgt_view : sqrtf(dxs * dxs + dys * dys + z * z) / den;
            float val;
            if (a.cone) {
                float vidx = (mag * z - a.v0) / a.det_h * a.adv / a.dv;
                float vff = floorf(vidx);
                if (vff < 0.f || vff + 1.f > vmax) continue;
                int vf = (int)vff;
                float rv = vidx - vff;
                size_t r0 = (size_t)vf * row_stride + vb;
                size_t r1 = r0 + row_stride;
                float s00 = sg[r0], s01 = sg[r0 + 1], s10 = sg[r1], s11 = sg[r1 + 1];
                if (a.mask) {
                    s00 *= a.mask[r0];
                    s01 *= a.mask[r0 + 1];
                    s10 *= a.mask[r1];
                    s11 *= a.mask[r1 + 1];
                }
                float p0 = s00 * (1.f - ru) + s01 * ru;
                float p1 = s10 * (1.f - ru) + s11 * ru;
                val = (1.f - rv) * p0 + rv * p1;
            } else {
                size_t r0 = (size_t)(z0 + k) * row_stride + vb;
                float s0 = sg[r0], s1 = sg[r0 + 1];
                if (a.mask) {
                    s0 *= a.mask[r0];
                    s1 *= a.mask[r0 + 1];
                }
                val = (1.f - ru) * s0 + ru * s1;
            }
            acc[k] += val * wgt * a.pix;
        }
    }
    const size_t plane = (size_t)a.nx * a.nx;
    float *out = a.vol + (size_t)iy * a.nx + ix;
#pragma unroll
    for (int k = 0; k < ZC; ++k)
        if (k < nk) out[(size_t)(z0 + k) * plane] = alive ? acc[k] : 0.f;
}

// Fan-beam IR back-projection from slice-chunk records.  The sinogram
// [slice][view][col] (times the ray mask, when one is given) is first
// repacked so that the ZC slices of a chunk sit together:
// rec[((c * nviews + j) * nu + col) * ZC + k] = sgm[c * ZC + k][j][col]
// (0 past the last slice).  A thread then owns ZC slices of one (x, y)
// column as in bp_ir_axial_kernel, and the two detector taps of all its
// slices in a view are the two adjacent records at col = uf and uf + 1:
// 2 * ZC * 4 contiguous bytes, read with 16-byte loads instead of 2 * ZC
// scalar ones.  Same arithmetic, in the same order, as bp_ir_axial_kernel
// (so the same results, bit for bit).
template <int ZC>
__global__ __launch_bounds__(256) void bp_ir_rec_pack_kernel(const float *__restrict__ sgm,
                                                             const float *__restrict__ mask, int nz,
                                                             int nviews, int nu,
                                                             float *__restrict__ rec) {
    // block: 256 / ZC consecutive columns of one (chunk, view), ZC slices each
    constexpr int kCols = 256 / ZC;
    __shared__ float tile[ZC][kCols + 1];
    const int nb = (nu + kCols - 1) / kCols;
    const int b = (int)blockIdx.x;
    const int cb = b % nb, t = b / nb, j = t % nviews, c = t / nviews;
    const int u0 = cb * kCols;
    const size_t row = (size_t)nviews * nu;
    // read: ZC rows of kCols columns (coalesced along the columns)
    {
        const int k = (int)threadIdx.x / kCols, uu = (int)threadIdx.x % kCols;
        const int z = c * ZC + k, col = u0 + uu;
        float v = 0.f;
        if (z < nz && col < nu) {
            const size_t i = (size_t)z * row + (size_t)j * nu + col;
            v = sgm[i];
            if (mask) v *= mask[i];
        }
        tile[k][uu] = v;
    }
    __syncthreads();
    // write: kCols records of ZC floats (contiguous)
    const int uu = (int)threadIdx.x / ZC, k = (int)threadIdx.x % ZC;
    const int col = u0 + uu;
    if (col < nu) rec[(((size_t)c * nviews + j) * nu + col) * ZC + k] = tile[k][uu];
}

template <int ZC, bool FLATZ>
__global__ __launch_bounds__(256) void bp_ir_rec_kernel(BpArgs a, const float *__restrict__ rec) {
    int ix, iy, zci;
    decode_tile(a, mct::xcd_remap((int)blockIdx.x, a.tiles_x * a.tiles_y * a.nzc, a.nxcd), ix, iy, zci);
    if (ix >= a.nx || iy >= a.nx) return;
    const int z0 = zci * ZC;
    const int nk = a.nz - z0 < ZC ? a.nz - z0 : ZC;
    float x, y;
    voxel_xy(a, ix, iy, x, y);
    const float halfz = (float)(a.nz - 1) / 2.0f;
    float zb[ZC], acc[ZC];
#pragma unroll
    for (int k = 0; k < ZC; ++k) {
        zb[k] = ((float)(z0 + k) - halfz) * a.vh;
        acc[k] = 0.f;
    }
    bool alive = true;
    const float4 *chunk = reinterpret_cast<const float4 *>(rec + (size_t)zci * a.nviews * a.nu * ZC);
    for (int j = 0; j < a.nviews; ++j) {
        const float4 t = a.vt[j];
        const float czv = a.center_z ? a.center_z[j] : a.cz;
        float L = a.sid - x * t.x - y * t.y;
        float dxs = x - a.sid * t.x;
        float dys = y - a.sid * t.y;
        bool dflag = fabsf(dxs) > fabsf(dys);
        float den = dflag ? fabsf(dxs) : fabsf(dys);
        float mag = a.sdd / L;
        float yr = (-x) * t.y + y * t.x;
        float u = a.curved ? a.sdd * atan2f(yr, L) : mag * yr;
        float uidx = (u - a.u0) / a.du;
        if (uidx < 0.f || uidx + 1.f > (float)(a.nu - 1)) {
            alive = false;
            break;
        }
        int uf = (int)floorf(uidx);
        uf = uf > a.nu - 2 ? a.nu - 2 : uf;
        float ru = uidx - (float)uf;
        // records uf and uf + 1 of view j: 2 * ZC floats
        const float4 *r = chunk + ((size_t)j * a.nu + uf) * (ZC / 4);
        float s[2 * ZC];
#pragma unroll
        for (int q = 0; q < ZC / 2; ++q) {
            const float4 v = r[q];
            s[4 * q] = v.x;
            s[4 * q + 1] = v.y;
            s[4 * q + 2] = v.z;
            s[4 * q + 3] = v.w;
        }
        const float wgt_view = FLATZ ? sqrtf(dxs * dxs + dys * dys + czv * czv) / den : 0.f;
#pragma unroll
        for (int k = 0; k < ZC; ++k) {
            const float z = zb[k] + czv;
            const float wgt = FLATZ ? wgt_view : sqrtf(dxs * dxs + dys * dys + z * z) / den;
            const float val = (1.f - ru) * s[k] + ru * s[ZC + k];
            acc[k] += val * wgt * a.pix;
        }
    }
    const size_t plane = (size_t)a.nx * a.nx;
    float *out = a.vol + (size_t)iy * a.nx + ix;
#pragma unroll
    for (int k = 0; k < ZC; ++k)
        if (k < nk) out[(size_t)(z0 + k) * plane] = alive ? acc[k] : 0.f;
}

// IR back-projection for every view mode and the P-matrix branch
// (run_mgfbp_iterative_recon_IRN_CGLS.py:384-474): one thread per voxel,
// same weight, tests and truncation flag as bp_ir_axial_kernel.
__global__ __launch_bounds__(256) void bp_ir_generic_kernel(BpArgs a) {
    const int64_t n = (int64_t)a.nz * a.nx * a.nx;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ix = (int)(id % a.nx);
    const int iy = (int)((id / a.nx) % a.nx);
    const int iz = (int)(id / ((int64_t)a.nx * a.nx));
    float x, y, zb;
    voxel_xyz(a, ix, iy, iz, 0.f, x, y, zb);
    const float *sg = reinterpret_cast<const float *>(a.sino);
    const size_t row_stride = (size_t)a.nviews * a.nu;
    const float vmax = (float)(a.nv - 1);
    float acc = 0.f;
    bool alive = true;
    for (int j = 0; j < a.nviews; ++j) {
        const float4 t = a.vt[j];
        const float z = zb + (a.center_z ? a.center_z[j] : a.cz);
        const float L = a.sid - x * t.x - y * t.y;
        const float dxs = x - a.sid * t.x;
        const float dys = y - a.sid * t.y;
        const float den = fabsf(dxs) > fabsf(dys) ? fabsf(dxs) : fabsf(dys);
        float mag, uidx;
        const float *P = a.apply_pm ? a.pm + 12 * j : nullptr;
        if (!a.apply_pm) {
            mag = a.sdd / L;
            const float yr = (-x) * t.y + y * t.x;
            const float u = a.curved ? a.sdd * atan2f(yr, L) : mag * yr;
            uidx = (u - a.u0) / a.du;
        } else {
            mag = 1.0f / (P[8] * x + P[9] * y + P[10] * z + P[11] * 1.0f);
            uidx = (P[0] * x + P[1] * y + P[2] * z + P[3] * 1.0f) * mag;
        }
        if (uidx < 0.f || uidx + 1.f > (float)(a.nu - 1)) {
            alive = false;
            break;
        }
        int uf = (int)floorf(uidx);
        uf = uf > a.nu - 2 ? a.nu - 2 : uf;
        const float ru = uidx - (float)uf;
        const float wgt = sqrtf(dxs * dxs + dys * dys + z * z) / den;
        const size_t vb = (size_t)j * a.nu + uf;
        float val;
        if (a.cone) {
            const float vidx = a.apply_pm ? (P[4] * x + P[5] * y + P[6] * z + P[7] * 1.0f) * mag
                                          : (mag * z - a.v0) / a.det_h * a.adv / a.dv;
            const float vff = floorf(vidx);
            if (vff < 0.f || vff + 1.f > vmax) continue;
            const int vf = (int)vff;
            const float rv = vidx - vff;
            const size_t r0 = (size_t)vf * row_stride + vb, r1 = r0 + row_stride;
            float s00 = sg[r0], s01 = sg[r0 + 1], s10 = sg[r1], s11 = sg[r1 + 1];
            if (a.mask) {
                s00 *= a.mask[r0];
                s01 *= a.mask[r0 + 1];
                s10 *= a.mask[r1];
                s11 *= a.mask[r1 + 1];
            }
            const float p0 = s00 * (1.f - ru) + s01 * ru;
            const float p1 = s10 * (1.f - ru) + s11 * ru;
            val = (1.f - rv) * p0 + rv * p1;
        } else {
            const size_t r0 = (size_t)iz * row_stride + vb;
            float s0 = sg[r0], s1 = sg[r0 + 1];
            if (a.mask) {
                s0 *= a.mask[r0];
                s1 *= a.mask[r0 + 1];
            }
            val = (1.f - ru) * s0 + ru * s1;
        }
        acc += val * wgt * a.pix;
    }
    a.vol[id] = alive ? acc : 0.f;
}

BpArgs make_args(const mct_plan *p) {
    const mct_geometry &g = p->g;
    BpArgs a = {};
    a.nx = g.img_dim;
    a.nz = g.img_dim_z;
    a.nu = g.n_u;
    a.nv = g.n_v;
    a.nviews = g.n_views;
    a.pix = g.pix;
    a.vh = g.vox_h;
    a.cx = g.cx;
    a.cy = g.cy;
    a.cz = g.cz;
    a.cos_rot = p->cos_rot;
    a.sin_rot = p->sin_rot;
    a.sid = g.sid;
    a.sdd = g.sdd;
    a.div_factor = p->div_factor;
    a.det_h = g.det_h;
    a.u0 = p->u0;
    a.du = p->du;
    a.v0 = p->v0;
    a.dv = p->dv;
    a.adv = p->adv;
    a.vinv = (g.cone && g.det_h > 0.f) ? (float)((double)p->adv / ((double)g.det_h * (double)p->dv)) : 0.f;
    a.inv_du = (float)(1.0 / (double)p->du);
    // largest float below n_v - 1: a clamped row coordinate floors to <= n_v - 2
    a.vlim = g.n_v >= 2 ? nextafterf((float)(g.n_v - 1), 0.f) : 0.f;
    a.ua = p->ua;
    a.ub = p->ub;
    a.da = p->da;
    a.db = p->db;
    a.curved = g.curved;
    a.cone = g.cone;
    a.view_mode = g.view_mode;
    a.apply_pm = g.apply_pmatrix;
    a.pm = p->pmatrix;
    a.center_z = p->center_z;
    a.vt = p->bp_view;
    a.nxcd = p->ctx->xcd_count;
    a.j0 = 0;
    a.nvb = g.n_views;
    a.accum = 0;
    a.zc0 = 0;
    return a;
}

#define MCT_AXIAL_LAUNCH(KERNEL, ZC, DT) \
    hipLaunchKernelGGL((KERNEL<ZC, DT>), dim3(nblocks), dim3(256), 0, ctx->stream, a)

// Window tables larger than this are not built (the direct kernel runs):
// 16 GiB, e.g. a 1024^3 volume from 7600 views (8 B per tile, 16-slice
// chunk and view; 2.3 GB at 1024^3 from 1080 views).
constexpr size_t kMaxWinTableBytes = size_t(16) << 30;
// Largest LDS window (cells) the staged kernel takes: 8 staging slots per
// thread; two FP32 buffers of 2048 cells fill 64 KiB.
constexpr int kLdsMaxCells = 2048;

// Builds the plan's column intervals and window table on first use; one
// synchronisation of the context's stream (to read the largest window).
int ensure_lds_tables(mct_plan *p, BpArgs a) {
    if (p->bp_win_state != 0) return p->bp_win_state > 0 ? MCT_OK : MCT_ERR_UNSUPPORTED;
    mct_ctx *ctx = p->ctx;
    p->bp_win_state = -1;
    const size_t entries = (size_t)a.tiles_x * a.tiles_y * a.nzc * a.nviews;
    if (entries * sizeof(BpWin) > kMaxWinTableBytes || a.tiles_x * a.tiles_y > 65535 ||
        a.nzc > 65535)
        return MCT_ERR_UNSUPPORTED;
    MCT_HIP(hipMalloc(&p->bp_iv, sizeof(int2) * (size_t)a.nx * a.nx));
    MCT_HIP(hipMalloc(&p->bp_win, sizeof(BpWin) * entries));
    int *maxes = nullptr;
    MCT_HIP(hipMalloc(&maxes, 4 * sizeof(int)));
    int host[4] = {0, 0, 0, 0};
    hipError_t e = hipMemsetAsync(maxes, 0, 4 * sizeof(int), ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bp_column_iv_kernel, dim3(mct::blocks_for(a.nx, 256), a.nx), dim3(256), 0,
                           ctx->stream, a, p->bp_iv, maxes + 2);
        hipLaunchKernelGGL(bp_window_kernel,
                           dim3(mct::blocks_for(a.nviews, 64), a.tiles_x * a.tiles_y, a.nzc), dim3(64),
                           0, ctx->stream, a, reinterpret_cast<BpWin *>(p->bp_win), maxes);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host, maxes, sizeof(host), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(maxes);
    if (e != hipSuccess)
        return mct::set_error(MCT_ERR_DEVICE, "back-projector tables: %s", hipGetErrorString(e));
    p->bp_win_w = host[0] > 0 ? host[0] : 1;
    p->bp_win_h = host[1] > 0 ? host[1] : 1;
    p->bp_zc_lo = host[3] > 0 ? (a.nz - host[2]) / kLdsZC : 0;
    p->bp_zc_hi = host[3] > 0 ? (host[3] - 1) / kLdsZC : -1;
    if ((int64_t)p->bp_win_w * p->bp_win_h > kLdsMaxCells) {
        hipFree(p->bp_iv);
        hipFree(p->bp_win);
        p->bp_iv = nullptr;
        p->bp_win = nullptr;
        return MCT_ERR_UNSUPPORTED;
    }
    p->bp_win_state = 1;
    return MCT_OK;
}

// Staging slots per thread for a window of `cells` packed cells (the
// instantiated NQ values) and the dynamic LDS of its two buffers.
inline int lds_slots(int cells) {
    const int nq = (cells + 255) / 256;
    return nq <= 2 ? 2 : nq <= 4 ? 4 : nq <= 6 ? 6 : 8;
}
inline size_t lds_bytes(int cells, int32_t dtype) {
    return 2 * (size_t)cells * (dtype == MCT_F32 ? 16 : 8);
}
// The instantiated NQ for a pitch x rows window read from `layout`: QUAD
// needs NQ * 256 staging slots >= the window's cells, ROWS NQ / 2 * 256
// strips of 4 cell rows >= pitch * ceil(rows / 4) (as few as that: each
// strip holds five element rows in registers); -1: none fits.
inline int lds_nq(int pitch, int rows, int32_t layout) {
    if (layout == MCT_ROWS) {
        const int64_t strips = (int64_t)pitch * ((rows + 3) / 4);
        const int64_t ns = (strips + 255) / 256;
        return ns <= 4 ? (int)(2 * ns) : -1;
    }
    return lds_slots(pitch * rows);
}
// Automatic selection declines windows whose two LDS buffers exceed this
// (occupancy would fall below 4 workgroups per CU).
constexpr size_t kLdsAutoBytes = 40 * 1024;

// Fills a.win / a.iv / a.pitch / a.rows (and the 16 x 16 x 16 tiling) for the
// LDS-staged kernel; MCT_ERR_UNSUPPORTED when the geometry does not allow it
// (the caller then runs the direct kernel or reports it).
int prepare_cone_lds(mct_plan *p, BpArgs &a, int32_t dtype, bool automatic) {
    const mct_geometry &g = p->g;
    if (!(g.cone && g.view_mode == 1 && !g.apply_pmatrix && a.vinv != 0.f && g.vox_h > 0.f &&
          g.n_u >= 2 && g.n_v >= 2))
        return MCT_ERR_UNSUPPORTED;
    a.tiles_x = (g.img_dim + kLdsTile - 1) / kLdsTile;
    a.tiles_y = a.tiles_x;
    a.nzc = (g.img_dim_z + kLdsZC - 1) / kLdsZC;
    int rc = ensure_lds_tables(p, a);
    if (rc) return rc;
    if (automatic && lds_bytes(p->bp_win_w * p->bp_win_h, dtype) > kLdsAutoBytes)
        return MCT_ERR_UNSUPPORTED;
    if (lds_nq(p->bp_win_w, p->bp_win_h, a.layout) < 0) return MCT_ERR_UNSUPPORTED;
    a.win = reinterpret_cast<const BpWin *>(p->bp_win);
    a.iv = p->bp_iv;
    a.pitch = p->bp_win_w;
    a.rows = p->bp_win_h;
    return MCT_OK;
}

int launch_cone_lds(mct_plan *p, BpArgs a, int32_t dtype, bool automatic) {
    int rc = prepare_cone_lds(p, a, dtype, automatic);
    if (rc) return rc;
    const int cells = a.pitch * a.rows;
    const size_t shmem = lds_bytes(cells, dtype);
    mct_ctx *ctx = p->ctx;
    // A whole-scan launch covers only the chunks where some column keeps a
    // voxel; the slices of the others are 0, written by a fill.  (Skipping
    // them also spreads the work evenly over the XCDs: at 512^2 x 256 / 1080
    // views the first and last of 16 chunks are empty, and the two XCDs that
    // held one of them had half the work of the others.)
    int nzc_run = a.nzc;
    if (!a.accum) {
        const int lo = p->bp_zc_lo, hi = p->bp_zc_hi < a.nzc - 1 ? p->bp_zc_hi : a.nzc - 1;
        nzc_run = hi >= lo ? hi - lo + 1 : 0;
        a.zc0 = hi >= lo ? lo : 0;
        const size_t plane = (size_t)a.nx * a.nx;
        const int64_t zlo = (int64_t)a.zc0 * kLdsZC;
        const int64_t zhi = nzc_run ? std::min<int64_t>((int64_t)(a.zc0 + nzc_run) * kLdsZC, a.nz) : 0;
        if (zlo > 0)
            hipLaunchKernelGGL(bp_fill_zero_kernel, dim3(mct::blocks_for((int64_t)(zlo * plane), 256)),
                               dim3(256), 0, ctx->stream, a.vol, (int64_t)(zlo * plane));
        if (zhi < a.nz)
            hipLaunchKernelGGL(bp_fill_zero_kernel, dim3(mct::blocks_for((int64_t)((a.nz - zhi) * plane), 256)),
                               dim3(256), 0, ctx->stream, a.vol + zhi * plane, (int64_t)((a.nz - zhi) * plane));
        if (nzc_run == 0) {
            MCT_LAUNCHED();
            return MCT_OK;
        }
    }
    const int nblocks = a.tiles_x * a.tiles_y * nzc_run;
#define LDS_LAUNCH2(DT, NQ, ROWS)                                                                 \
    do {                                                                                              \
        if (a.accum)                                                                                  \
            hipLaunchKernelGGL((bp_cone_lds_kernel<DT, NQ, true, ROWS>), dim3(nblocks), dim3(256),    \
                               shmem, ctx->stream, a);                                                \
        else                                                                                          \
            hipLaunchKernelGGL((bp_cone_lds_kernel<DT, NQ, false, ROWS>), dim3(nblocks), dim3(256),   \
                               shmem, ctx->stream, a);                                                \
    } while (0)
#define LDS_LAUNCH(DT, NQ)                                    \
    do {                                                      \
        if (a.layout == MCT_ROWS) LDS_LAUNCH2(DT, NQ, true);  \
        else LDS_LAUNCH2(DT, NQ, false);                      \
    } while (0)
#define LDS_DT(NQ)                                         \
    if (dtype == MCT_F32) LDS_LAUNCH(MCT_F32, NQ);         \
    else if (dtype == MCT_F16) LDS_LAUNCH(MCT_F16, NQ);    \
    else LDS_LAUNCH(MCT_BF16, NQ);
    switch (lds_nq(a.pitch, a.rows, a.layout)) {
        case 2: LDS_DT(2) break;
        case 4: LDS_DT(4) break;
        case 6: LDS_DT(6) break;
        default: LDS_DT(8) break;
    }
#undef LDS_DT
#undef LDS_LAUNCH
#undef LDS_LAUNCH2
    MCT_LAUNCHED();
    return MCT_OK;
}

// Largest number of jobs in one batched launch.
constexpr int kMaxBatchJobs = 4096;

// Validates the jobs and fills the job table: args[i] and offs[i] (first
// work item of job i; offs[njobs] = total).  Cone jobs get the LDS-staged
// tile where the context allows it (building their plans' tables on first
// use); lds_cells is the largest LDS window among them (0: none).
int fill_batch(mct_ctx *ctx, const mct_bp_job *jobs, int njobs, int dtype, BpArgs *args, int *offs,
               int &total_out, int &lds_cells, int &lds_q) {
    MCT_CHECK(dtype == MCT_F32 || dtype == MCT_F16 || dtype == MCT_BF16, "bad dtype %d", dtype);
    int64_t total = 0;
    lds_cells = 0;
    lds_q = 0;
    for (int i = 0; i < njobs; ++i) {
        const mct_bp_job &jb = jobs[i];
        MCT_CHECK(jb.plan && jb.filt && jb.vol, "job %d: NULL plan, sinogram or volume", i);
        MCT_CHECK(jb.plan->ctx == ctx, "job %d: plan belongs to another context", i);
        MCT_CHECK(jb.dtype == dtype, "job %d: all jobs of a batch must share one dtype", i);
        const mct_geometry &g = jb.plan->g;
        MCT_CHECK(g.view_mode == 1 && !g.apply_pmatrix && !g.curved,
                  "job %d: batched back-projection supports flat-detector axial jobs", i);
        MCT_CHECK(g.cone ? (jb.layout == MCT_QUAD || jb.layout == MCT_ROWS) : jb.layout == MCT_PAIR,
                  "job %d: cone reads QUAD or ROWS, fan PAIR", i);
        if (!g.cone) MCT_CHECK(g.n_v == g.img_dim_z, "job %d: fan beam needs one row per slice", i);
        if (g.cone) MCT_CHECK(g.det_h > 0.f, "job %d: cone beam needs det_h > 0", i);
        BpArgs a = make_args(jb.plan);
        a.sino = jb.filt;
        a.layout = jb.layout;
        a.vol = jb.vol;
        const int zc = g.cone ? 16 : 8;
        a.tiles_x = (g.img_dim + 15) / 16;
        a.tiles_y = a.tiles_x;
        a.nzc = (g.img_dim_z + zc - 1) / zc;
        if (g.cone && ctx->bp_kernel != MCT_BP_DIRECT) {
            const bool automatic = ctx->bp_kernel == MCT_BP_AUTO;
            int rc = prepare_cone_lds(jb.plan, a, dtype, automatic);
            if (rc == MCT_OK) {
                const int cells = a.pitch * a.rows, nq = lds_nq(a.pitch, a.rows, a.layout);
                lds_cells = cells > lds_cells ? cells : lds_cells;
                lds_q = nq > lds_q ? nq : lds_q;
            } else if (rc != MCT_ERR_UNSUPPORTED) {
                return rc;
            } else if (!automatic) {
                return mct::set_error(MCT_ERR_UNSUPPORTED,
                                      "job %d: LDS-staged back-projection does not fit its geometry", i);
            }
        }
        args[i] = a;
        offs[i] = (int)total;
        total += (int64_t)a.tiles_x * a.tiles_y * a.nzc;
        MCT_CHECK(total < (int64_t)1 << 30, "batch too large");
    }
    offs[njobs] = (int)total;
    total_out = (int)total;
    return MCT_OK;
}

// The persistent batch kernel over a device job table, on the context's
// stream: as many workgroups as stay resident together.  Launches only (no
// copy, allocation or synchronisation), so it can be captured in a graph.
int launch_batch(mct_ctx *ctx, const BpArgs *dargs, const int *doffs, int njobs, int total, int dtype,
                 int lds_cells, int lds_q) {
    const size_t shmem = lds_cells ? lds_bytes(lds_cells, dtype) : 0;
    const int nq = lds_cells ? lds_q : 0;
#define BATCH_KERNEL(DT, NQ) bp_batch_kernel<DT, NQ>
#define BATCH_LAUNCH(DT, NQ)                                                                        \
    do {                                                                                            \
        int per_cu = 0;                                                                             \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, BATCH_KERNEL(DT, NQ), 256, shmem) \
                != hipSuccess || per_cu < 1)                                                        \
            per_cu = 1;                                                                             \
        int64_t grid = (int64_t)ctx->cu_count * per_cu;                                             \
        grid = grid < total ? grid : total;                                                         \
        hipLaunchKernelGGL(BATCH_KERNEL(DT, NQ), dim3((unsigned)grid), dim3(256), shmem,            \
                           ctx->stream, dargs, doffs, njobs, total);                                \
    } while (0)
#define BATCH_DT(NQ)                                         \
    if (dtype == MCT_F32) BATCH_LAUNCH(MCT_F32, NQ);         \
    else if (dtype == MCT_F16) BATCH_LAUNCH(MCT_F16, NQ);    \
    else BATCH_LAUNCH(MCT_BF16, NQ);
    switch (nq) {
        case 0: BATCH_DT(0) break;
        case 2: BATCH_DT(2) break;
        case 4: BATCH_DT(4) break;
        case 6: BATCH_DT(6) break;
        default: BATCH_DT(8) break;
    }
#undef BATCH_DT
#undef BATCH_LAUNCH
#undef BATCH_KERNEL
    MCT_LAUNCHED();
    return MCT_OK;
}

// FBP back-projection of views [j0, j0 + nvb) with volume mode accum (see
// BpArgs); mct_backproject_fbp is the whole scan with accum 0.
int backproject(mct_plan *plan, const void *filt, int32_t dtype, int32_t layout, int j0, int nvb,
                int accum, float *vol) {
    MCT_CHECK(plan && filt && vol, "NULL argument");
    const mct_geometry &g = plan->g;
    mct_ctx *ctx = plan->ctx;
    MCT_CHECK(dtype == MCT_F32 || dtype == MCT_F16 || dtype == MCT_BF16, "bad dtype %d", dtype);
    if (g.cone)
        MCT_CHECK(layout == MCT_QUAD || layout == MCT_ROWS || layout == MCT_PLAIN,
                  "cone-beam back-projection reads QUAD, ROWS or PLAIN taps (got %d)", layout);
    else
        MCT_CHECK(layout == MCT_PAIR || layout == MCT_PLAIN,
                  "fan-beam back-projection reads PAIR or PLAIN taps (got %d)", layout);
    if (!g.cone)
        MCT_CHECK(g.n_v == g.img_dim_z,
                  "fan beam: one detector row per slice required (rows %d, slices %d)", g.n_v,
                  g.img_dim_z);
    if (g.cone && !g.apply_pmatrix) MCT_CHECK(g.det_h > 0.f, "cone beam needs det_h > 0");
    BpArgs a = make_args(plan);
    a.sino = filt;
    a.layout = layout;
    a.vol = vol;
    a.j0 = j0;
    a.nvb = nvb;
    a.accum = accum;
    const bool fast = g.view_mode == 1 && !g.apply_pmatrix && layout != MCT_PLAIN;
    if (g.cone && (layout == MCT_QUAD || layout == MCT_ROWS) && ctx->bp_kernel != MCT_BP_DIRECT) {
        const bool automatic = ctx->bp_kernel == MCT_BP_AUTO;
        int rc = launch_cone_lds(plan, a, dtype, automatic);
        if (rc == MCT_OK) {
            ctx->bp_last = MCT_BP_LDS;
            return MCT_OK;
        }
        if (rc != MCT_ERR_UNSUPPORTED) return rc;
        if (!automatic)
            return mct::set_error(MCT_ERR_UNSUPPORTED,
                                  "LDS-staged back-projection does not fit this geometry");
    }
    if (fast) {
        int zc = ctx->bp_zc ? ctx->bp_zc : (g.cone ? 16 : 8);
        a.tiles_x = (g.img_dim + 15) / 16;
        a.tiles_y = a.tiles_x;
        a.nzc = (g.img_dim_z + zc - 1) / zc;
        int nblocks = a.tiles_x * a.tiles_y * a.nzc;
        ctx->bp_last = g.cone ? MCT_BP_DIRECT : 3;
        if (g.cone) {
            int bt = ctx->bp_batch ? ctx->bp_batch : 16;
            bt = bt > zc ? zc : bt;
#define CONE_L(ZC, B, L)                                                                          \
    if (dtype == MCT_F32)                                                                         \
        hipLaunchKernelGGL((bp_cone_axial_kernel<ZC, B, MCT_F32, L>), dim3(nblocks), dim3(256), 0,  \
                           ctx->stream, a);                                                       \
    else if (dtype == MCT_F16)                                                                    \
        hipLaunchKernelGGL((bp_cone_axial_kernel<ZC, B, MCT_F16, L>), dim3(nblocks), dim3(256), 0,  \
                           ctx->stream, a);                                                       \
    else                                                                                          \
        hipLaunchKernelGGL((bp_cone_axial_kernel<ZC, B, MCT_BF16, L>), dim3(nblocks), dim3(256), 0, \
                           ctx->stream, a);
#define CONE(ZC, B)                        \
    if (layout == MCT_ROWS) { CONE_L(ZC, B, true) } \
    else { CONE_L(ZC, B, false) }
            if (zc == 4) { CONE(4, 4) }
            else if (zc == 8 && bt == 4) { CONE(8, 4) }
            else if (zc == 8) { CONE(8, 8) }
            else if (zc == 16 && bt == 4) { CONE(16, 4) }
            else if (zc == 16 && bt == 8) { CONE(16, 8) }
            else if (zc == 16) { CONE(16, 16) }
            else if (bt == 4) { CONE(32, 4) }
            else if (bt == 8) { CONE(32, 8) }
            else { CONE(32, 16) }
#undef CONE
#undef CONE_L
        } else {
#define CASE_DT(ZC)                                                                  \
    if (dtype == MCT_F32) MCT_AXIAL_LAUNCH(bp_fan_axial_kernel, ZC, MCT_F32);        \
    else if (dtype == MCT_F16) MCT_AXIAL_LAUNCH(bp_fan_axial_kernel, ZC, MCT_F16);   \
    else MCT_AXIAL_LAUNCH(bp_fan_axial_kernel, ZC, MCT_BF16);
            switch (zc) {
                case 4: CASE_DT(4) break;
                case 8: CASE_DT(8) break;
                case 16: CASE_DT(16) break;
                default: CASE_DT(32) break;
            }
#undef CASE_DT
        }
    } else {
        int64_t n = (int64_t)g.img_dim_z * g.img_dim * g.img_dim;
        unsigned nblocks = mct::blocks_for(n, 256);
        ctx->bp_last = 4;
        if (dtype == MCT_F32)
            hipLaunchKernelGGL(bp_generic_kernel<MCT_F32>, dim3(nblocks), dim3(256), 0, ctx->stream, a);
        else if (dtype == MCT_F16)
            hipLaunchKernelGGL(bp_generic_kernel<MCT_F16>, dim3(nblocks), dim3(256), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL(bp_generic_kernel<MCT_BF16>, dim3(nblocks), dim3(256), 0, ctx->stream, a);
    }
    MCT_LAUNCHED();
    return MCT_OK;
}

}  // namespace

extern "C" {

int mct_backproject_fbp(mct_plan *plan, const void *filt, int32_t dtype, int32_t layout, float *vol) {
    MCT_CHECK(plan, "NULL argument");
    return backproject(plan, filt, dtype, layout, 0, plan->g.n_views, 0, vol);
}

int mct_backproject_views(mct_plan *plan, const void *filt, int32_t dtype, int32_t layout,
                          int32_t view_begin, int32_t view_count, int32_t finish, float *vol) {
    MCT_CHECK(plan, "NULL argument");
    MCT_CHECK(view_begin >= 0 && view_count >= 1 && view_count <= plan->g.n_views - view_begin,
              "views [%d, %d + %d) outside the scan's %d views", view_begin, view_begin, view_count,
              plan->g.n_views);
    return backproject(plan, filt, dtype, layout, view_begin, view_count, finish ? 2 : 1, vol);
}

int mct_backproject_ir(mct_plan *plan, const float *sgm, const float *ray_mask, float *vol) {
    MCT_CHECK(plan && sgm && vol, "NULL argument");
    const mct_geometry &g = plan->g;
    mct_ctx *ctx = plan->ctx;
    if (!g.cone)
        MCT_CHECK(g.n_v == g.img_dim_z, "fan beam: one detector row per slice required");
    else if (!g.apply_pmatrix)
        MCT_CHECK(g.det_h > 0.f, "cone beam needs det_h > 0");
    BpArgs a = make_args(plan);
    a.sino = sgm;
    a.mask = ray_mask;
    a.layout = MCT_PLAIN;
    a.vol = vol;
    if (g.view_mode != 1 || g.apply_pmatrix) {
        const int64_t n = (int64_t)g.img_dim_z * g.img_dim * g.img_dim;
        hipLaunchKernelGGL(bp_ir_generic_kernel, dim3(mct::blocks_for(n, 256)), dim3(256), 0,
                           ctx->stream, a);
        MCT_LAUNCHED();
        return MCT_OK;
    }
    a.tiles_x = (g.img_dim + 15) / 16;
    a.tiles_y = a.tiles_x;
    if (!g.cone && g.img_dim_z >= 8) {
        // fan beam, 8 slices or more: slice-chunk records (bp_ir_rec_kernel;
        // 8 slices per thread measured faster than 4, 16 or 32 at 64 x 512^2,
        // 120 views x 800 columns)
        constexpr int zc = 8;
        a.nzc = (g.img_dim_z + zc - 1) / zc;
        const int64_t nrec = (int64_t)a.nzc * g.n_views * g.n_u * zc;
        int rc = mct::ensure_buffer(ctx, &ctx->bp_rec, &ctx->bp_rec_cap, nrec);
        if (rc) return rc;
        const int nb = (g.n_u + 256 / zc - 1) / (256 / zc);
        hipLaunchKernelGGL(bp_ir_rec_pack_kernel<zc>, dim3(nb * g.n_views * a.nzc), dim3(256), 0,
                           ctx->stream, sgm, ray_mask, g.img_dim_z, g.n_views, g.n_u, ctx->bp_rec);
        MCT_LAUNCHED();
        const int nblocks = a.tiles_x * a.tiles_y * a.nzc;
        if (g.vox_h == 0.f)
            hipLaunchKernelGGL((bp_ir_rec_kernel<zc, true>), dim3(nblocks), dim3(256), 0, ctx->stream,
                               a, ctx->bp_rec);
        else
            hipLaunchKernelGGL((bp_ir_rec_kernel<zc, false>), dim3(nblocks), dim3(256), 0, ctx->stream,
                               a, ctx->bp_rec);
        MCT_LAUNCHED();
        return MCT_OK;
    }
    const int zc = 8;
    a.nzc = (g.img_dim_z + zc - 1) / zc;
    int nblocks = a.tiles_x * a.tiles_y * a.nzc;
    if (!g.cone && g.vox_h == 0.f)
        hipLaunchKernelGGL((bp_ir_axial_kernel<8, true>), dim3(nblocks), dim3(256), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((bp_ir_axial_kernel<8, false>), dim3(nblocks), dim3(256), 0, ctx->stream, a);
    MCT_LAUNCHED();
    return MCT_OK;
}

int mct_backproject_batch(mct_ctx *ctx, const mct_bp_job *jobs, int32_t njobs) {
    MCT_CHECK(ctx && jobs, "NULL argument");
    MCT_CHECK(njobs >= 0 && njobs <= kMaxBatchJobs, "njobs must be in [0, %d] (got %d)", kMaxBatchJobs,
              njobs);
    if (njobs == 0) return MCT_OK;
    // every call rewrites a staging slot that a captured copy would read
    MCT_CHECK(!ctx->frozen, "context is frozen: mct_backproject_batch rewrites its staging slots; "
                            "captured graphs use mct_bp_batch_create / mct_bp_batch_launch");
    const int dtype = jobs[0].dtype;
    const size_t args_bytes = sizeof(BpArgs) * (size_t)njobs;
    const size_t bytes = args_bytes + sizeof(int) * ((size_t)njobs + 1);
    // A staging slot (pinned table + device copy) whose previous batch has
    // finished: its event follows that batch's kernel.  With every slot still
    // in flight, wait for the oldest; the host does not otherwise wait for the
    // stream, so a scheduler can queue other work behind a long batch.
    int slot = -1;
    for (int i = 0; i < mct::kBatchSlots && slot < 0; ++i) {
        const int k = (ctx->batch_next + i) % mct::kBatchSlots;
        mct::BatchSlot &b = ctx->batch_slot[k];
        if (!b.ev) {
            slot = k;
            break;
        }
        const hipError_t q = hipEventQuery(b.ev);
        if (q == hipSuccess)
            slot = k;
        else if (q != hipErrorNotReady)
            return mct::set_error(MCT_ERR_DEVICE, "batch staging event: %s", hipGetErrorString(q));
    }
    if (slot < 0) {
        slot = ctx->batch_next;
        MCT_HIP(hipEventSynchronize(ctx->batch_slot[slot].ev));
    }
    ctx->batch_next = (slot + 1) % mct::kBatchSlots;
    mct::BatchSlot &b = ctx->batch_slot[slot];
    if (!b.ev) MCT_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
    if (b.cap < (int64_t)bytes) {  // the slot's last batch has finished: free to replace
        if (b.host) hipHostFree(b.host);
        if (b.dev) hipFree(b.dev);
        b.host = b.dev = nullptr;
        b.cap = 0;
        const size_t cap = bytes < 65536 ? 65536 : bytes;
        MCT_HIP(hipHostMalloc(&b.host, cap, hipHostMallocDefault));
        MCT_HIP(hipMalloc(&b.dev, cap));
        b.cap = (int64_t)cap;
    }
    BpArgs *args = reinterpret_cast<BpArgs *>(b.host);
    int *offs = reinterpret_cast<int *>(reinterpret_cast<char *>(b.host) + args_bytes);
    int total = 0, lds_cells = 0, lds_q = 0;
    int rc = fill_batch(ctx, jobs, njobs, dtype, args, offs, total, lds_cells, lds_q);
    if (rc) return rc;
    MCT_HIP(hipMemcpyAsync(b.dev, b.host, bytes, hipMemcpyHostToDevice, ctx->stream));
    const BpArgs *dargs = reinterpret_cast<const BpArgs *>(b.dev);
    const int *doffs = reinterpret_cast<const int *>(reinterpret_cast<const char *>(b.dev) + args_bytes);
    rc = launch_batch(ctx, dargs, doffs, njobs, total, dtype, lds_cells, lds_q);
    const hipError_t e = hipEventRecord(b.ev, ctx->stream);  // after the kernel: slot reusable
    if (rc) return rc;
    if (e != hipSuccess) return mct::set_error(MCT_ERR_DEVICE, "batch staging event: %s", hipGetErrorString(e));
    return MCT_OK;
}

int mct_bp_batch_create(mct_ctx *ctx, const mct_bp_job *jobs, int32_t njobs, mct_bp_batch **out) {
    MCT_CHECK(out, "out is NULL");
    *out = nullptr;
    MCT_CHECK(ctx && jobs, "NULL argument");
    MCT_CHECK(njobs >= 1 && njobs <= kMaxBatchJobs, "njobs must be in [1, %d] (got %d)", kMaxBatchJobs,
              njobs);
    const int dtype = jobs[0].dtype;
    std::vector<BpArgs> args((size_t)njobs);
    std::vector<int> offs((size_t)njobs + 1);
    int total = 0, lds_cells = 0, lds_q = 0;
    int rc = fill_batch(ctx, jobs, njobs, dtype, args.data(), offs.data(), total, lds_cells, lds_q);
    if (rc) return rc;
    const size_t args_bytes = sizeof(BpArgs) * (size_t)njobs;
    void *dev = nullptr;
    MCT_HIP(hipMalloc(&dev, args_bytes + sizeof(int) * offs.size()));
    hipError_t e = hipMemcpy(dev, args.data(), args_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(reinterpret_cast<char *>(dev) + args_bytes, offs.data(), sizeof(int) * offs.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(dev);
        return mct::set_error(MCT_ERR_DEVICE, "batch table upload: %s", hipGetErrorString(e));
    }
    mct_bp_batch *b = new mct_bp_batch();
    b->ctx = ctx;
    b->device = ctx->device;
    b->dev = dev;
    b->args_bytes = args_bytes;
    b->njobs = njobs;
    b->total = total;
    b->dtype = dtype;
    b->lds_cells = lds_cells;
    b->lds_q = lds_q;
    *out = b;
    return MCT_OK;
}

int mct_bp_batch_launch(mct_bp_batch *b) {
    MCT_CHECK(b, "NULL batch");
    const BpArgs *dargs = reinterpret_cast<const BpArgs *>(b->dev);
    const int *doffs = reinterpret_cast<const int *>(reinterpret_cast<const char *>(b->dev) + b->args_bytes);
    return launch_batch(b->ctx, dargs, doffs, b->njobs, b->total, b->dtype, b->lds_cells, b->lds_q);
}

int mct_bp_batch_destroy(mct_bp_batch *b) {
    if (!b) return MCT_OK;
    hipSetDevice(b->device);
    if (b->dev) hipFree(b->dev);
    delete b;
    return MCT_OK;
}

}  // extern "C"
