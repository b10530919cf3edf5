/*
 * bp_oracle.c - C restatement of Mgfbp.BackProjectionPixelDriven
 * (run_mgfbp.py:1062-1187), axial view, no P-matrix, flat detector.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ (as a second, independent
 * restatement beside oracle/recon.py) and by bench.py's cpu_baseline leg
 * (timed on the host cores).  Never linked into the product.
 *
 * Loop nesting follows the Taichi kernel: parallel over the (x, y) columns
 * (the outer ndrange, here an OpenMP loop), serial over slices and views.
 * FP32 in the reference's expression order; build without FMA contraction
 * (-ffp-contract=off).  cos/sin of each view come from the caller (FP64
 * values rounded once to FP32, as the library's geometry plan does).
 */
#include <math.h>
#include <stdint.h>

#define PI_REF 3.1415926536f

/* filt: [nrows][nviews][nu]; vol: [z1 - z0][nx][nx] (slices z0..z1-1 of nz). */
int bp_oracle_fbp(const float *filt, int nrows, int nviews, int nu, const float *cosv,
                  const float *sinv, const float *dbeta, const float *u, const float *v, int nx,
                  int nz, int z0, int z1, float pix, float vh, float cx, float cy, float cz,
                  float sid, float sdd, float det_h, float div_factor, int cone, float *vol) {
    const float half = (float)(nx - 1) / 2.0f;
    const float halfz = (float)(nz - 1) / 2.0f;
    const float u0 = u[0], du = u[1] - u[0];
    const float v0 = v[0];
    const float dv = nrows > 1 ? v[1] - v[0] : 1.0f;
    const float adv = fabsf(dv);
    const int64_t plane = (int64_t)nx * nx;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t col = 0; col < plane; ++col) {
        const int ix = (int)(col % nx), iy = (int)(col / nx);
        const float x = pix * ((float)ix - half) + cx;
        const float y = (-pix) * ((float)iy - half) + cy;
        for (int iz = z0; iz < z1; ++iz) {
            const float z = ((float)iz - halfz) * vh + cz;
            float acc = 0.0f;
            int alive = 1;
            for (int j = 0; j < nviews; ++j) {
                const float L = sid - x * cosv[j] - y * sinv[j];
                const float mag = sdd / L;
                const float yr = (-x) * sinv[j] + y * cosv[j];
                const float uu = mag * yr;
                const float uidx = (uu - u0) / du;
                if (uidx < 0.0f || uidx + 1.0f > (float)(nu - 1)) {
                    alive = 0;
                    break;
                }
                const int uf = (int)floorf(uidx);
                const float ru = uidx - (float)uf;
                const float w = 1.0f / (L * L);
                float val;
                if (cone) {
                    const float vidx = (mag * z - v0) / det_h * adv / dv;
                    const float vff = floorf(vidx);
                    if (vff < 0.0f || vff + 1.0f > (float)(nrows - 1)) {
                        alive = 0;
                        break;
                    }
                    const int vf = (int)vff;
                    const float rv = vidx - vff;
                    const float *r0 = filt + ((int64_t)vf * nviews + j) * nu;
                    const float *r1 = r0 + (int64_t)nviews * nu;
                    const float p0 = r0[uf] * (1.0f - ru) + r0[uf + 1] * ru;
                    const float p1 = r1[uf] * (1.0f - ru) + r1[uf + 1] * ru;
                    val = (1.0f - rv) * p0 + rv * p1;
                } else {
                    const float *r0 = filt + ((int64_t)iz * nviews + j) * nu;
                    val = (1.0f - ru) * r0[uf] + ru * r0[uf + 1];
                }
                acc = acc + (sid * w) * val * dbeta[j] * div_factor;
            }
            vol[(int64_t)(iz - z0) * plane + col] = alive ? acc : 0.0f;
        }
    }
    return 0;
}

/*
 * Horizontal pass of Mgfbp.ConvolveSgmAndKernel (run_mgfbp.py:1033-1046),
 * FP32: out[r][k] = (sum over m ascending of x[r][m] * taps[k + n-1 - m]) *
 * scale, O(n^2) per row as in the Taichi kernel, parallel over the rows.
 * x, out: [rows][n]; taps: 2n - 1 entries.
 */
int filter_oracle_f32(const float *x, int64_t rows, int n, const float *taps, float scale,
                      float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; ++r) {
        const float *xr = x + r * n;
        float *o = out + r * n;
        for (int k = 0; k < n; ++k) {
            float acc = 0.0f;
            for (int m = 0; m < n; ++m) acc += xr[m] * taps[k + n - 1 - m];
            o[k] = acc * scale;
        }
    }
    return 0;
}

/* OpenMP thread count of the restatements (the CPU baseline's 1-core and
 * all-core legs). */
#include <omp.h>
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }
