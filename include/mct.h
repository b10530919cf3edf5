/*
 * mct.h - C ABI of the MI355X (gfx950) CT reconstruction hot path.
 *
 * Plain pointers and sizes only.  "DEVICE" arguments are GPU memory the
 * caller owns (the Python layer allocates them as torch ROCm tensors); the
 * library never frees caller memory.  "HOST" arguments are read during the
 * call and not kept.  Every kernel is queued asynchronously on the stream
 * given to mct_set_stream (default: the null stream); mct_synchronize waits.
 * Every function returns MCT_OK (0) or a negative mct_status; the message of
 * the last failure on the calling thread is mct_last_error().  No entry point
 * aborts or exits the process.
 *
 * The reference (SEU-CT-Recon/Reconstruction_program_taichi, "MandoCT-Taichi")
 * has no native code: its device code is Taichi @ti.kernel methods that the
 * Python classes Mgfbp / Mgfpj / Mgfbp_ir call.  Each declaration below names
 * the Taichi kernel or Python method (file:line) whose behaviour it replaces.
 */
#ifndef MCT_H
#define MCT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCT_ABI_VERSION 11

typedef enum mct_status {
    MCT_OK = 0,
    MCT_ERR_INVALID = -1,     /* bad argument or geometry */
    MCT_ERR_DEVICE = -2,      /* HIP runtime error */
    MCT_ERR_UNSUPPORTED = -3, /* option combination not implemented */
    MCT_ERR_NOMEM = -4        /* allocation failed or a fixed capacity exceeded */
} mct_status;

/* element type of a sinogram buffer read by the back-projector */
typedef enum mct_dtype { MCT_F32 = 0, MCT_F16 = 1, MCT_BF16 = 2 } mct_dtype;

/* memory layout of a filtered sinogram read by the back-projector
 *   PLAIN: [rows][views][n_u] scalars (the reference's field order)
 *   QUAD : [views][rows+1][n_u-1] x 4 taps {(r,c),(r,c+1),(r+1,c),(r+1,c+1)}:
 *          the four bilinear taps of detector cell (r,c), r = -1 .. rows-1,
 *          in one vector load (rows -1 and `rows` read as zeros)
 *   PAIR : [rows][views][n_u-1] x 2 taps {(r,c),(r,c+1)} (fan beam)
 *   ROWS : [views][rows+2][n_u-1] x {(r,c), (r,c+1) - (r,c)}, r = -1 .. rows
 *          (rows -1 and `rows` read as zeros): cone beam, half the bytes of
 *          QUAD; the back-projector forms a cell's four taps from two rows */
typedef enum mct_layout { MCT_PLAIN = 0, MCT_QUAD = 1, MCT_PAIR = 2, MCT_ROWS = 3 } mct_layout;

/* ramp-filter arithmetic */
typedef enum mct_filter_method {
    MCT_FILTER_F32_DIRECT = 0, /* FP32 VALU direct convolution, reference summation order */
    MCT_FILTER_BF16_MFMA = 1,  /* Toeplitz GEMM on MFMA, BF16 operands, FP32 accumulate */
    MCT_FILTER_BF16X3_MFMA = 2 /* Toeplitz GEMM on MFMA, split-BF16 (hi*hi+hi*lo+lo*hi) */
} mct_filter_method;

/*
 * Scan and image geometry: the derived members of Mgfbp.__init__
 * (run_mgfbp.py:87-719), Mgfpj.__init__ (run_mgfpj.py:98-449) and
 * Mgfpj_v3.__init__ (run_mgfpj_v3.py:129-181).  The arrays are HOST FP32
 * arrays produced in the reference's order of operations
 * (GenerateAngleArray run_mgfbp.py:930-934, GenerateDectPixPosArray
 * run_mgfbp.py:937-944 and run_mgfpj.py:457-461); mct_plan_create copies them
 * to the device.
 */
typedef struct mct_geometry {
    /* image grid */
    int32_t img_dim;     /* ImageDimension (x and y) */
    int32_t img_dim_z;   /* slices (fan beam: one per detector row) */
    float pix;           /* PixelSize [mm] */
    float vox_h;         /* VoxelHeight [mm] (fan beam: 0) */
    float cx, cy, cz;    /* ImageCenter x, y and ImageCenterZ [mm] */
    float img_rot;       /* ImageRotation [rad] (added into angles already) */
    int32_t view_mode;   /* ReconViewMode: 1 axial, 2 coronal, 3 sagittal */
    /* scan */
    int32_t n_views;
    float sid, sdd;          /* SourceIsocenterDistance, SourceDetectorDistance [mm] */
    float total_scan_angle;  /* [rad] */
    int32_t short_scan;
    int32_t cone;            /* ConeBeam */
    int32_t curved;          /* CurvedDetector */
    /* detector */
    int32_t n_u;         /* elements per row (forward projection: count x OversampleSize) */
    int32_t n_v;         /* rows described by v[] */
    int32_t n_v_full;    /* DetectorElementCountVertical (helical pitch) */
    float det_h;         /* DetectorElementHeight [mm] */
    /* forward projection */
    float fpj_step;      /* ForwardProjectionStepSize */
    int32_t fpj_count_mode; /* steps = 0: floor((l_max - l_min)/(step*diag))
                               (run_mgfpj.py:526-527); 1: floor(D/(step*diag))
                               (run_mgfpj_v3.py:252) */
    int32_t helical;
    float helical_pitch;
    int32_t apply_pmatrix;
    /* HOST arrays */
    const float *angles;   /* [n_views] rad (ImageRotation / StartAngle included) */
    const float *u;        /* [n_u] detector u positions [mm] */
    const float *v;        /* [n_v] detector v positions [mm] */
    const float *pmatrix;  /* [12*n_views] or NULL */
    const float *center_z; /* [n_views] per-view ImageCenterZ or NULL (use cz) */
    const float *matrix_a; /* [9*n_views] forward-projection P-matrix branch or NULL */
    const float *x_s;      /* [3*n_views] source positions of that branch or NULL */
} mct_geometry;

typedef struct mct_ctx mct_ctx;
typedef struct mct_plan mct_plan;

/* ---- context ------------------------------------------------------------ */
int mct_create(int device, mct_ctx **out);
int mct_destroy(mct_ctx *ctx);
/* hip_stream: a hipStream_t (NULL = null stream) */
int mct_set_stream(mct_ctx *ctx, void *hip_stream);
int mct_synchronize(mct_ctx *ctx);

/* Views [view_begin, view_begin + view_count) of a HOST FP32 sinogram
 * [n_v][n_views][n_u] into a DEVICE block [n_v][view_count][n_u]: one
 * two-dimensional asynchronous copy on `stream` (a hipStream_t; NULL: the
 * context's stream).  For the copy to overlap kernels, host_sgm must be
 * pinned (hipHostMalloc) memory.  Replaces the per-view upload
 * img_sgm_taichi.from_numpy(img_sgm[:, view_idx:view_idx+1, :]) of
 * run_mgfbp_v2.py:71. */
int mct_copy_views(mct_ctx *ctx, void *stream, const float *host_sgm, int32_t n_v, int32_t n_views,
                   int32_t n_u, int32_t view_begin, int32_t view_count, float *dev_block);
const char *mct_last_error(void);
int mct_abi_version(void);

/* ---- geometry plan -------------------------------------------------------
 * Replaces the Taichi fields that Mgfbp.__init__ allocates and fills
 * (run_mgfbp.py:675-687, InitializeArrays :1244-1259): uploads the arrays of
 * *g and the per-view tables derived from them (cos/sin of the view angle with
 * and without ImageRotation, the angular interval of run_mgfbp.py:1105-1111).
 * cos/sin are evaluated in double precision and rounded to FP32, so every
 * kernel sees the same table.  The plan keeps no pointer into *g.  Operations
 * on a plan run on its context's stream; mct_plan_destroy does not use the
 * context, so plans and contexts may be destroyed in either order. */
int mct_plan_create(mct_ctx *ctx, const mct_geometry *g, mct_plan **out);
int mct_plan_destroy(mct_plan *plan);

/* ---- FBP / FDK ----------------------------------------------------------- */

/* Mgfbp.BHCorrection run_mgfbp.py:946-954 (when bh_order > 0) followed by
 * Mgfbp.WeightSgm run_mgfbp.py:956-1011: p <- sum_t c_t p^(t+1), then the
 * cosine weight (flat or curved detector) and, for short scans, the Parker
 * weight.  In place on a DEVICE FP32 sinogram [n_v][n_views][n_u].
 * bh_coef: HOST array of bh_order floats (bh_order <= 16). */
int mct_weight_sinogram(mct_plan *plan, float *sgm, const float *bh_coef, int32_t bh_order);

/* The same weighting for a block of views, as the view-by-view FBP
 * Mgfbp_v2.BHCorrection / Mgfbp_v2.WeightSgm run_mgfbp_v2.py:120-186 applies
 * it: sgm is a DEVICE FP32 [n_v][view_count][n_u] block holding views
 * [view_begin, view_begin + view_count) of the plan's scan; the Parker weight
 * uses each view's angle in the whole scan. */
int mct_weight_views(mct_plan *plan, float *sgm, int32_t view_begin, int32_t view_count,
                     const float *bh_coef, int32_t bh_order);

/* Vertical Gaussian of Mgfbp.ConvolveSgmAndKernel run_mgfbp.py:1019-1031:
 * out[i,j,k] = det_h * sum_{|i-n|<=10} in[n,j,k] * kern[i + rows-1 - n].
 * in/out DEVICE [rows][views][n_u]; kern HOST array of 2*rows-1 taps. */
int mct_filter_vertical(mct_ctx *ctx, const float *in, float *out, int32_t rows,
                        int32_t views, int32_t n_u, const float *kern, float det_h);

/* Ramp row convolution of Mgfbp.ConvolveSgmAndKernel run_mgfbp.py:1033-1046:
 * out[r,k] = scale * sum_m in[r,m] * taps[k + n-1 - m] for each of n_rows rows
 * of length n: the full, non-circular linear convolution, i.e. a Toeplitz
 * product.  in/out DEVICE FP32 [n_rows][n]; taps HOST array of 2n-1 floats
 * (GenerateHammingKernel run_mgfbp.py:877-914 or its Gaussian-apodised form).
 * method: mct_filter_method.  For the MFMA methods the split taps are
 * cached in the context and rebuilt when taps or n change, and the input rows
 * are first converted into BF16 planes held in context scratch (2 bytes per
 * padded sample, 4 for the split method; it grows to the largest call), so
 * calls on one context must be ordered, as calls on its stream are. */
int mct_filter_rows(mct_ctx *ctx, const float *in, float *out, int64_t n_rows, int32_t n,
                    const float *taps, float scale, int32_t method);

/* BHCorrection + WeightSgm (run_mgfbp.py:946-1011) and the ramp convolution
 * (run_mgfbp.py:1033-1046) of views [view_begin, view_begin + view_count) in
 * one pass over the sinogram: the same as mct_weight_views on sgm followed
 * by mct_filter_rows(plan's context, sgm, out, n_v * view_count, n_u, ...),
 * bit for bit, except that sgm is left unweighted (the weighting is applied
 * where the MFMA methods convert their input rows to BF16 planes).
 * sgm DEVICE FP32 [n_v][view_count][n_u]; out DEVICE FP32, same shape;
 * method: MCT_FILTER_BF16_MFMA or MCT_FILTER_BF16X3_MFMA only.  There is no
 * vertical Gaussian between the two steps here. */
int mct_weight_filter_rows(mct_plan *plan, const float *sgm, float *out, int32_t view_begin,
                           int32_t view_count, const float *bh_coef, int32_t bh_order,
                           const float *taps, float scale, int32_t method);

/* mct_weight_filter_rows followed by mct_pack_sinogram(..., MCT_F32,
 * MCT_ROWS), bit for bit, without the FP32 filtered sinogram: the filter's
 * output is written straight as the back-projector's ROWS taps
 * (run_mgfbp.py:946-1046 and the filtered field img_sgm_filtered_taichi,
 * :1230-1242).  packed DEVICE, mct_packed_bytes(n_v, view_count, n_u,
 * MCT_F32, MCT_ROWS) bytes; needs n_v >= 2 and n_u - 1 not a multiple of 64
 * (else filter and pack separately). */
int mct_weight_filter_pack(mct_plan *plan, const float *sgm, void *packed, int32_t view_begin,
                           int32_t view_count, const float *bh_coef, int32_t bh_order,
                           const float *taps, float scale, int32_t method);

/* Repack a DEVICE FP32 filtered sinogram [rows][views][n_u] into the dtype and
 * layout the back-projector reads.  Replaces the reference's Taichi field
 * img_sgm_filtered_taichi (run_mgfbp.py:1230-1242). */
int mct_pack_sinogram(mct_ctx *ctx, const float *in, int32_t rows, int32_t views, int32_t n_u,
                      void *out, int32_t dtype, int32_t layout);

/* Bytes of a packed buffer (host-only helper; -1 for an invalid shape). */
int64_t mct_packed_bytes(int32_t rows, int32_t views, int32_t n_u, int32_t dtype,
                         int32_t layout);

/* Mgfbp.BackProjectionPixelDriven run_mgfbp.py:1062-1187: pixel-driven
 * FBP/FDK back-projection with the view weight sid/L^2 * dbeta * div_factor;
 * a voxel that projects off the detector in any view is 0.
 * filt: DEVICE packed sinogram (QUAD for cone, PAIR for fan, or PLAIN).
 * vol: DEVICE FP32 [img_dim_z][img_dim][img_dim], overwritten. */
int mct_backproject_fbp(mct_plan *plan, const void *filt, int32_t dtype, int32_t layout,
                        float *vol);

/* View-by-view FBP, Mgfbp_v2.MainFunction run_mgfbp_v2.py:62-94 with
 * Mgfbp_v2.BackProjectionPixelDriven :238-364 and SetTruncatedRegionToZero
 * :232-236, for a block of views at a time: back-projects views
 * [view_begin, view_begin + view_count) of the plan's scan from filt, a
 * packed block holding only those views (as mct_pack_sinogram writes it for
 * view_count views), and ADDS them to vol, voxel by voxel in view order, so a
 * scan streamed block after block into a zeroed volume sums every voxel in
 * the same order as mct_backproject_fbp.  A voxel that projects off the
 * detector is set to -10000 (the reference's truncation mark; later blocks
 * keep adding to it).  finish != 0 (the last block): every voxel below -5000
 * is then set to 0.  The view weight dbeta uses the neighbouring views of
 * the whole scan. */
int mct_backproject_views(mct_plan *plan, const void *filt, int32_t dtype, int32_t layout,
                          int32_t view_begin, int32_t view_count, int32_t finish, float *vol);

/* One job of a batched back-projection: a flat-detector axial plan, its
 * packed filtered sinogram (QUAD for cone beam, PAIR for fan beam) and the
 * output volume. */
typedef struct mct_bp_job {
    mct_plan *plan;
    const void *filt;
    int32_t dtype;
    int32_t layout;
    float *vol;
} mct_bp_job;

/* Mgfbp.BackProjectionPixelDriven run_mgfbp.py:1062-1187 for njobs volumes or
 * slice stacks of different geometry and size in ONE persistent launch (the
 * reference back-projects one file per kernel call, run_mgfbp.py:58-81).
 * Each job's result equals mct_backproject_fbp's for that job alone.  All
 * jobs share one dtype; plans must come from ctx.  The job table is copied
 * to the device through one of eight context-owned pinned staging slots; the
 * call waits on the host only when all eight still hold batches in flight. */
int mct_backproject_batch(mct_ctx *ctx, const mct_bp_job *jobs, int32_t njobs);

/* The same batched back-projection, prepared once and launched many times:
 * mct_bp_batch_create validates the jobs, builds the plans' tables and copies
 * the job table to device memory (synchronously); mct_bp_batch_launch then
 * only launches the persistent kernel on the context's current stream - no
 * copy, allocation or synchronisation - so it can be captured into a HIP
 * graph and replayed.  Every replay back-projects the same sinogram buffers
 * into the same volumes (whatever they hold at that time).  Plans and buffers
 * must outlive the batch. */
typedef struct mct_bp_batch mct_bp_batch;
int mct_bp_batch_create(mct_ctx *ctx, const mct_bp_job *jobs, int32_t njobs, mct_bp_batch **out);
int mct_bp_batch_launch(mct_bp_batch *batch);
int mct_bp_batch_destroy(mct_bp_batch *batch);

/* Tile shape of the axial back-projector (autotuner hook): zc slices per
 * thread (4, 8, 16 or 32) and, for cone beam, batch taps loads issued
 * together (4, 8 or 16, at most zc); 0 restores a default.  Applies to the
 * direct-gather kernel. */
int mct_set_bp_tile(mct_ctx *ctx, int32_t zc, int32_t batch);

/* Cone-beam axial back-projector used by mct_backproject_fbp (QUAD taps):
 *   MCT_BP_AUTO   the LDS-staged kernel where the geometry allows it
 *                 (flat or curved detector, no P-matrix, detector windows
 *                 that fit in LDS), else the direct-gather kernel;
 *   MCT_BP_DIRECT every voxel-view update gathers its taps from memory;
 *   MCT_BP_LDS    each 16 x 16-column x 16-slice tile stages the detector
 *                 window it reads in a view into LDS once (MCT_ERR_UNSUPPORTED
 *                 where the geometry does not allow it).
 * The LDS-staged kernel builds two geometry tables in the plan on its first
 * use (one synchronisation of the context's stream). */
typedef enum mct_bp_kernel { MCT_BP_AUTO = 0, MCT_BP_DIRECT = 1, MCT_BP_LDS = 2 } mct_bp_kernel;
int mct_set_bp_kernel(mct_ctx *ctx, int32_t kind);
/* Kernel the last mct_backproject_fbp call on ctx launched: MCT_BP_DIRECT,
 * MCT_BP_LDS, 3 (fan-beam axial), 4 (generic: other view modes, P-matrix,
 * PLAIN taps) or 0 (none yet); -1 for a NULL ctx. */
int mct_bp_last_kernel(const mct_ctx *ctx);

/* ---- forward projection -------------------------------------------------- */

/* Mgfpj.ForwardProjectionBilinear run_mgfpj.py:463-607 followed by
 * Mgfpj.BinSinogram run_mgfpj.py:609-616, for every detector row and view in
 * one launch; also the per-view variants Mgfpj_v3.ForwardProjectionBilinear
 * run_mgfpj_v3.py:184-366 and Mgfbp_ir.ForwardProjectionBilinear
 * run_mgfbp_iterative_recon_IRN_CGLS.py:494-676 (ray mask).
 * vol DEVICE [img_dim_z][img_dim][img_dim]; out DEVICE
 * [n_v][n_views][n_u/oversample] (row r = detector row r of v[]).
 * Fan beam (cone == 0) reads slice r for detector row r.
 * ray_mask: NULL or DEVICE [n_v][n_views][n_u/oversample]; rays whose mask is
 * not 1 are left 0 (oversample must be 1).  oversample: OversampleSize.
 * apply_pmatrix: the P-matrix branch of run_mgfpj_v2.py:334-355 (sample l of
 * ray (u index i, row r, view j) at (l/|d - s|) A_j [i, r, 1] + x_s_j, rotated
 * by angles[0]); matrix_a and x_s are then required. */
int mct_forward_project(mct_plan *plan, const float *vol, float *out, const float *ray_mask,
                        int32_t oversample);

/* Forward projector used by mct_forward_project:
 *   MCT_FP_AUTO    the cell kernel where each order of cells is under 4 GiB,
 *                  else the direct one;
 *   MCT_FP_DIRECT  every sample gathers its 8 (cone) or 4 (fan) voxels from
 *                  the volume as given;
 *   MCT_FP_CELLS   the volume is first repacked, in context-owned scratch,
 *                  into cells of bilinear coefficients (cone: planes z and
 *                  z+1, 32 B; fan: 16 B) in an x-major and a y-major order;
 *                  each view reads the order whose fast axis follows its
 *                  detector rows, so a sample is one or two coalesced 16-byte
 *                  loads (MCT_ERR_UNSUPPORTED where it does not fit).
 * Both take the same samples with the same in-bounds tests; the cells hold
 * FP32 bilinear coefficients, so the sums agree to rounding. */
typedef enum mct_fp_kernel { MCT_FP_AUTO = 0, MCT_FP_DIRECT = 1, MCT_FP_CELLS = 2 } mct_fp_kernel;
int mct_set_fp_kernel(mct_ctx *ctx, int32_t kind);
/* Kernel the last mct_forward_project call on ctx ran (MCT_FP_DIRECT or
 * MCT_FP_CELLS; 0 none yet; -1 for a NULL ctx). */
int mct_fp_last_kernel(const mct_ctx *ctx);
/* Cell projector: when a launch has fewer wavefronts than waves_per_cu per CU,
 * each ray's steps are split over up to 16 wavefronts (partial sums added in a
 * fixed order by a second pass) to reach it.  0 restores the default (256).
 * The split changes the order of summation: results with different settings
 * agree to rounding, results with the same setting are identical. */
int mct_set_fp_split(mct_ctx *ctx, int32_t waves_per_cu);

/* No counterpart in the reference: for callers that capture the library's
 * launches in a HIP graph.  A captured graph holds the addresses of the
 * context's scratch (filter planes and tap panels, the forward projector's
 * packed volume and partial sums, the IR back-projector's records, reduction
 * partials) and the contents of its cached tables (ramp taps, vertical
 * kernel, helical positions).  With frozen = 1, any call that would grow,
 * reallocate or rewrite one of them fails with MCT_ERR_INVALID instead
 * (calls of the captured sizes and taps still run); 0 lifts it. */
int mct_set_frozen(mct_ctx *ctx, int32_t frozen);

/* Mgfbp_ir.GenEffectiveMapForwardProjectionAgent
 * run_mgfbp_iterative_recon_IRN_CGLS.py:679-833: 1 for rays whose samples
 * inside the in-plane FOV disc all stay within the volume's z range, else 0
 * (cone beam; fan beam gives all ones); P-matrix rays as in
 * mct_forward_project.  mask DEVICE [n_v][n_views][n_u]. */
int mct_ray_mask(mct_plan *plan, float *mask);

/* Mgfpj.AddPossionNoise run_mgfpj.py:643-654: N = I0 e^-p; N += g sqrt(N)
 * with g standard normal; N <= 0 -> 1e-6; p = ln(I0 / N).  g comes from a
 * counter-based generator (Philox4x32-10, Box-Muller) keyed by seed, so the
 * result depends only on (seed, element index).  In place, DEVICE. */
int mct_add_noise(mct_ctx *ctx, float *sgm, int64_t n, float photon_number, uint64_t seed);

/* ---- helical rebinning ------------------------------------------------------ */

/* Scalars of Mgfbp_helical.SinogramRebinning run_mgfbp_helical.py:165-198
 * (FP32, as the reference passes them). */
typedef struct mct_helical {
    int32_t n_u;                /* DetectorElementCountHorizontal */
    int32_t n_rows;             /* detector rows of the helical sinogram */
    int32_t n_views;            /* views of the helical sinogram */
    int32_t view_count;         /* views of each rebinned fan sinogram */
    float det_offset_vertical;  /* DetectorOffsetVertical [mm] */
    float det_h;                /* DetectorElementHeight [mm] */
    float sdd;                  /* SourceDetectorDistance [mm] */
    float mag;                  /* sdd / sid */
    float dis_per_view;         /* table feed per view [mm] */
} mct_helical;

/* Mgfbp_helical.SinogramRebinning for nz slice positions in ONE launch (the
 * reference runs one launch per slice, run_mgfbp_helical.py:121-126):
 * out[z][i][j] = linear interpolation between the detector rows around
 * v_pos = (z_pos[z] - k dz) mag - offset in view k = floor(k0 - view_count/2
 * + 1/2) + i, k0 = (z_pos[z] - offset/mag)/dz, scaled by
 * sdd/sqrt(sdd^2 + v_pos^2); 0 where k < 0 or k > n_views - 2.
 * sgm DEVICE FP32 [n_rows][n_views][n_u]; z_pos HOST [nz];
 * out DEVICE FP32 [nz][view_count][n_u]. */
int mct_helical_rebin(mct_ctx *ctx, const mct_helical *h, const float *sgm, const float *z_pos,
                      int32_t nz, float *out);

/* ---- iterative reconstruction --------------------------------------------- */

/* Mgfbp_ir.BackProjectionPixelDrivenPerView
 * run_mgfbp_iterative_recon_IRN_CGLS.py:372-474 over ALL views in one launch,
 * then SetTruncatedRegionToZero :477-481: unfiltered back-projection with the
 * ray-length weight; a voxel whose u falls off the detector in any view is 0.
 * Every ReconViewMode and the P-matrix branch (:423-458).
 * sgm DEVICE FP32 [n_v][n_views][n_u]; ray_mask NULL or the same shape
 * (multiplies sgm); vol DEVICE [img_dim_z][img_dim][img_dim], overwritten. */
int mct_backproject_ir(mct_plan *plan, const float *sgm, const float *ray_mask, float *vol);

/* FP64 inner products, each written to one DEVICE double; deterministic (fixed
 * reduction tree independent of the stream's other work).
 * mct_dot_f64: out[0] = sum a[i]*b[i]. */
int mct_dot_f64(mct_ctx *ctx, const float *a, const float *b, int64_t n, double *out);

/* mct_dot_f64 without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving or iteration loops, harnesses that forbid
 * hipMalloc).  Same kernels and result as mct_dot_f64 (bit for bit); its 1024
 * FP64 partial sums live in a static 8 KiB device array instead of the
 * context's lazily allocated buffer, so any n fits.  Every call on one
 * device shares that array: calls that overlap on one device (several
 * streams or contexts) race and must be serialised by the caller. */
int mct_dot_f64_noalloc(mct_ctx *ctx, const float *a, const float *b, int64_t n, double *out);

/* Mgfbp_ir.GenerateWR run_mgfbp_iterative_recon_IRN_CGLS.py:234-243:
 * wr = 1/sqrt(ux^2 + uy^2 + uz^2 + beta) of y = x - prior (prior may be NULL),
 * with forward differences along axis 1 (ux), axis 2 (uy) and axis 0 (uz,
 * only when couple_z != 0 and nz > 2 - the reference couples slices exactly
 * then; couple_z = 0 treats the slices as independent images); the last
 * index of each axis has difference 0.
 * Evaluated in FP64, stored FP32.  All arrays DEVICE [nz][ny][nx]. */
int mct_tv_weight(mct_ctx *ctx, const float *x, const float *prior, int32_t nz, int32_t ny,
                  int32_t nx, int32_t couple_z, double beta, float *wr);

/* Mgfbp_ir.Dt_W_D run_mgfbp_iterative_recon_IRN_CGLS.py:267-286 scaled:
 * out = base + scale * Dt W D x (base may be NULL = 0; may alias out;
 * couple_z as for mct_tv_weight). */
int mct_dtwd(mct_ctx *ctx, const float *x, const float *wr, int32_t nz, int32_t ny, int32_t nx,
             int32_t couple_z, float scale, const float *base, float *out);

/* out = a*x + b*y elementwise, FP32 (a, b host scalars). */
int mct_axpby(mct_ctx *ctx, float a, const float *x, float b, const float *y, float *out,
              int64_t n);

/* One CG step of Mgfbp_ir.TikhonovSol
 * run_mgfbp_iterative_recon_IRN_CGLS.py:198-217 with its scalars on the
 * device in FP64.  s: DEVICE double[4]; on entry s[0] = <r,r>, s[1] = <d,q>.
 * mct_cg_update:    alpha = s[0]/s[1]; x += alpha d; r -= alpha q;
 *                   s[2] = <r,r> (new).
 * mct_cg_direction: beta = s[2]/s[0]; d = r + beta d; s[0] = s[2]. */
int mct_cg_update(mct_ctx *ctx, float *x, float *r, const float *d, const float *q, double *s,
                  int64_t n);
int mct_cg_direction(mct_ctx *ctx, float *d, const float *r, double *s, int64_t n);

/* Terms of Mgfbp_ir.LossValCalc run_mgfbp_iterative_recon_IRN_CGLS.py:184-186
 * in FP64 into one DEVICE double:
 * mct_sq_diff_f64: sum (a[i] - b[i]*mb[i])^2 (mb NULL = 1);
 * mct_tv_value:    TVPenaltyVal :245-254, sum sqrt(ux^2+uy^2+uz^2+beta) of
 *                  x - prior (prior may be NULL). */
int mct_sq_diff_f64(mct_ctx *ctx, const float *a, const float *b, const float *mb, int64_t n,
                    double *out);
int mct_tv_value(mct_ctx *ctx, const float *x, const float *prior, int32_t nz, int32_t ny,
                 int32_t nx, int32_t couple_z, double beta, double *out);

#ifdef __cplusplus
}
#endif
#endif /* MCT_H */
