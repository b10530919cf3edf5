// Internal declarations shared by the HIP translation units of libmct.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/mct.h"

namespace mct {
// A job-table staging slot of mct_backproject_batch: pinned host table, its
// device copy, and an event recorded after the batch's kernel (the slot is
// free again once it has completed).
struct BatchSlot {
    void *host = nullptr, *dev = nullptr;
    int64_t cap = 0;
    hipEvent_t ev = nullptr;
};
constexpr int kBatchSlots = 8;
}  // namespace mct

struct mct_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    int xcd_count = 8;
    int bp_zc = 0;     // slices per thread in the axial back-projector (0: default)
    int bp_batch = 0;  // loads issued together by the cone back-projector (0: default)
    int bp_kernel = 0; // cone back-projector: 0 automatic, 1 direct gather, 2 LDS-staged
    int bp_last = 0;   // kernel of the last mct_backproject_fbp (mct_bp_last_kernel)
    int fp_kernel = 0; // forward projector: 0 automatic, 1 direct taps, 2 packed cells
    int fp_last = 0;   // kernel of the last mct_forward_project (mct_fp_last_kernel)
    int fp_split = 0;  // cell projector: target wavefronts per CU when splitting rays (0: default)
    int frozen = 0;    // 1: scratch and cached tables may not grow or change (mct_set_frozen)
    // packed volume of the forward projector (two orders of bilinear cells)
    float *fp_pack = nullptr;
    int64_t fp_pack_cap = 0;  // floats
    float *fp_partial = nullptr;  // per-segment partial sums of split rays
    int64_t fp_partial_cap = 0;
    // fan IR back-projector: the (masked) sinogram in slice-chunk records
    float *bp_rec = nullptr;
    int64_t bp_rec_cap = 0;
    // partial sums of the deterministic FP64 reductions
    double *partials = nullptr;
    int64_t partials_cap = 0;
    // Toeplitz panel cache of mct_filter_rows (MFMA methods)
    void *toep = nullptr;
    int64_t toep_bytes = 0;
    int toep_n = 0, toep_method = -1;
    float toep_scale = 0.f;
    uint64_t toep_hash = 0;
    // split-BF16 planes (hi, lo) of the MFMA filter's input rows
    void *fsplit = nullptr;
    int64_t fsplit_bytes = 0;
    float *taps_dev = nullptr;  // device copy of the ramp taps (direct method)
    int64_t taps_cap = 0;
    uint64_t taps_hash = 0;
    float *vkern_dev = nullptr;  // device copy of the vertical Gaussian taps
    int64_t vkern_cap = 0;
    uint64_t vkern_hash = 0;
    float *aux_dev = nullptr;  // device copy of a small host table (helical slice positions)
    int64_t aux_cap = 0;
    uint64_t aux_hash = 0;
    // job tables of mct_backproject_batch: a ring of staging slots
    mct::BatchSlot batch_slot[mct::kBatchSlots];
    int batch_next = 0;
};

struct mct_plan {
    mct_ctx *ctx = nullptr;  // used by the operations, never by mct_plan_destroy
    int device = 0;
    mct_geometry g{};  // scalars; the host pointers are cleared
    // device copies of the geometry arrays
    float *angles = nullptr, *u = nullptr, *v = nullptr;
    float *pmatrix = nullptr, *center_z = nullptr, *matrix_a = nullptr, *x_s = nullptr;
    // per-view tables
    float4 *bp_view = nullptr;  // {cos(b_j - rot), sin(b_j - rot), dbeta_j, sid*dbeta_j*div}
    float2 *fp_view = nullptr;  // {cos b_j, sin b_j}
    // host-side scalars derived once
    float u0 = 0.f, du = 1.f;           // u[0], u[1] - u[0]
    float v0 = 0.f, dv = 1.f, adv = 1.f;  // v[0], v[1] - v[0], |v[1] - v[0]|
    float cos_rot = 1.f, sin_rot = 0.f;
    float div_factor = 1.f;
    float cos_a0 = 1.f, sin_a0 = 0.f;  // of angles[0] (P-matrix forward projection)
    float zpv = 0.f;                   // helical z shift per view (forward projection)
    int fp_count_mode = 0;
    // Sample distances of the cell forward projector, l[k] = k * step * diag
    // + l_min in FP32 (built on its first use) and the scalars they came from
    float *fp_ltab = nullptr;
    int fp_ltab_n = 0;
    float fp_ltab_key[3] = {0.f, 0.f, 0.f};
    // Exact detector-range tests of run_mgfbp.py:1142-1144 and 1170-1173 as
    // intervals: the u test passes iff ua <= u <= ub for u = fl(mag * yr);
    // the v test passes iff da <= d <= db for d = fl(fl(mag * z) - v[0]).
    float ua = 0.f, ub = 0.f, da = 0.f, db = 0.f;
    // Tables of the LDS-staged cone back-projector, built on its first use
    // (geometry only): per image column the exact interval of slices that
    // pass every view's range tests, and per (16 x 16-column tile, 16-slice
    // chunk, view) the window of packed detector cells the tile can read.
    int2 *bp_iv = nullptr;
    void *bp_win = nullptr;
    int bp_win_w = 0, bp_win_h = 0;  // largest window over all entries
    int bp_win_state = 0;            // 0 not built, 1 usable, -1 unusable geometry
    // chunks holding a voxel that some column keeps (every voxel outside
    // [bp_zc_lo, bp_zc_hi] is 0 in a whole-scan FBP; hi < lo: none)
    int bp_zc_lo = 0, bp_zc_hi = -1;
};

// A prepared batched back-projection (mct_bp_batch_create): the job table in
// device memory and the launch shape.
struct mct_bp_batch {
    mct_ctx *ctx = nullptr;
    int device = 0;
    void *dev = nullptr;  // BpArgs[njobs] followed by int offsets[njobs + 1]
    size_t args_bytes = 0;
    int njobs = 0, total = 0, dtype = 0, lds_cells = 0, lds_q = 0;
};

namespace mct {

constexpr float kPI = 3.1415926536f;  // the reference's PI (run_mgfbp.py:16)

// Word 3 of a raw buffer resource on gfx9-family GPUs (gfx950 included):
// 32-bit untyped data, no swizzle.  Loads beyond num_records return 0.
constexpr int kBufferDword3 = 0x00020000;

// Records the message returned by mct_last_error() for the calling thread.
int set_error(int code, const char *fmt, ...);
// Error for a frozen context's scratch or cached table that would change
// from `have` to `need` (mct_set_frozen).
int refuse_frozen(const char *what, int64_t have, int64_t need);

#define MCT_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return ::mct::set_error(MCT_ERR_DEVICE, "%s failed: %s", #call,     \
                                    hipGetErrorString(e_));                     \
    } while (0)

#define MCT_CHECK(cond, ...)                                                    \
    do {                                                                        \
        if (!(cond)) return ::mct::set_error(MCT_ERR_INVALID, __VA_ARGS__);     \
    } while (0)

#define MCT_LAUNCHED()                                                          \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess)                                                   \
            return ::mct::set_error(MCT_ERR_DEVICE, "kernel launch: %s",        \
                                    hipGetErrorString(e_));                     \
    } while (0)

int ensure_partials(mct_ctx *ctx, int64_t n);
// Takes the device address of mct_dot_f64_noalloc's static scratch on the
// current device (once per device; mct_create calls it).
int noalloc_scratch_init(int device);

// FNV-1a hash of a byte range (tap-set identity for the upload caches).
inline uint64_t fnv1a(const void *p, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// Copies n host floats to a context-owned device buffer unless the buffer
// already holds exactly these values (same hash and count): a repeated call
// then queues no copy, so the operation that uses the buffer only launches
// kernels and can be captured into a graph once it has run.
int upload_cached(mct_ctx *ctx, float **buf, int64_t *cap, uint64_t *hash, const float *host,
                  int64_t n);

// Grows a context-owned FP32 device buffer to at least n elements (waits for
// the stream before freeing the old one).
int ensure_buffer(mct_ctx *ctx, float **buf, int64_t *cap, int64_t n);

// Helical z shift per view of run_mgfpj.py:518-523 (0 without a helical scan).
float helical_z_per_view(const mct_plan *p, const float *angles_host);

// div_factor of run_mgfbp.py:1072-1084 evaluated in FP32.
float div_factor_of(const mct_geometry &g);

inline unsigned blocks_for(int64_t n, int threads) {
    return (unsigned)((n + threads - 1) / threads);
}

// Workgroups are dealt round robin over the XCDs; this maps the hardware
// workgroup id so that XCD k receives the contiguous logical range
// [k*per, (k+1)*per) and its L2 only sees that range's footprint.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks, int nxcd) {
    if (nxcd <= 1 || nblocks % nxcd != 0) return bid;
    int per = nblocks / nxcd;
    return (bid % nxcd) * per + bid / nxcd;
}

// Round-to-nearest-even FP32 -> BF16 / FP16 bit patterns and back.
__device__ __forceinline__ uint16_t f32_to_bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f32_to_f16_rn(float f) {
    _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}

}  // namespace mct
