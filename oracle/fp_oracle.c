/*
 * fp_oracle.c - C restatements of the forward projector
 * (Mgfpj.ForwardProjectionBilinear, run_mgfpj.py:463-607) and the IR
 * back-projector (Mgfbp_ir.BackProjectionPixelDrivenPerView over all views
 * plus SetTruncatedRegionToZero, run_mgfbp_iterative_recon_IRN_CGLS.py:
 * 372-481), axial view, no P-matrix, each in FP32 and FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the same arithmetic as oracle/recon.py's
 * forward_project / backproject_ir in the same expression order, so that
 * large cases (full-size fan scans, a whole cone detector row band) finish
 * in seconds; pinned against the NumPy restatement by tests/test_oracle.py.
 * Also timed by bench.py's cpu_baseline leg.  Never linked into the
 * product.  Build without FMA contraction (-ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>

/* ----------------------------------------------------------------------
 * Forward projection.  vol: [nz][nx][nx]; u: nuo entries (already
 * oversampled); v: nrows entries; out: [nrows][nviews][nuo].  center_z:
 * per-view image centre (NULL: cz); zpv: helical z shift per view;
 * count_mode 1: the v3 step count (floor(D / step)).
 * -------------------------------------------------------------------- */
#define DEFINE_FP(NAME, T)                                                                        \
    int NAME(const T *vol, int nx, int nz, T pix, T vh, T cx, T cy, T cz, const T *center_z,      \
             T sid, T sdd, const T *ang, int nviews, const T *u, int nuo, const T *v, int nrows,  \
             int cone, int curved, T step, int count_mode, T zpv, T *out) {                       \
        const T dim = (T)nx * pix;                                                                \
        const T dimz = (T)nz * vh;                                                                \
        const T D = (T)sqrt((T)2 * (dim * dim) + dimz * dimz);                                   \
        const T lmin = sid - D / (T)2;                                                            \
        const T lmax = sid + D / (T)2;                                                            \
        const T diag = (T)sqrt((T)2 * (pix * pix) + vh * vh);                                    \
        const T span = count_mode == 1 ? D : (lmax - lmin);                                       \
        const int64_t count = (int64_t)floor(span / (step * diag));                              \
        const T x0 = -((T)nx - (T)1) / (T)2 * pix + cx;                                          \
        const T y0 = -((T)nx - (T)1) / (T)2 * (-pix) + cy;                                       \
        const T zbase = -((T)nz - (T)1) / (T)2 * vh;                                             \
        const int64_t plane = (int64_t)nx * nx;                                                   \
        const int64_t nrays = (int64_t)nrows * nviews * nuo;                                      \
        _Pragma("omp parallel for schedule(dynamic, 256)")                                        \
        for (int64_t ray = 0; ray < nrays; ++ray) {                                               \
            const int i = (int)(ray % nuo);                                                       \
            const int j = (int)((ray / nuo) % nviews);                                            \
            const int r = (int)(ray / ((int64_t)nuo * nviews));                                   \
            T dx, dy;                                                                             \
            if (curved) {                                                                         \
                const T gp = (-u[i]) / sdd;                                                       \
                dx = (-sdd) * (T)cos((double)gp) + sid;                                           \
                dy = (-sdd) * (T)sin((double)gp);                                                 \
            } else {                                                                              \
                dx = -(sdd - sid);                                                                \
                dy = u[i];                                                                        \
            }                                                                                     \
            const T zs = zpv * (T)j;                                                              \
            const T dz = v[r] + zs;                                                               \
            const T ex = dx - sid, ey = dy - (T)0, ez = dz - zs;                                  \
            const T dist = (T)sqrt(ex * ex + ey * ey + ez * ez);                                  \
            const T ux = ex / dist, uy = ey / dist, uz = ez / dist;                               \
            const T ca = (T)cos((double)ang[j]), sa = (T)sin((double)ang[j]);                     \
            const T z0 = zbase + (center_z ? center_z[j] : cz);                                   \
            T acc = (T)0;                                                                         \
            for (int64_t k = 0; k < count; ++k) {                                                 \
                const T l = (T)k * step * diag + lmin;                                            \
                const T x = sid + ux * l;                                                         \
                const T y = (T)0 + uy * l;                                                        \
                const T z = zs + uz * l;                                                          \
                const T xr = x * ca - y * sa;                                                     \
                const T yr = y * ca + x * sa;                                                     \
                const T fxi = (T)floor((xr - x0) / pix);                                          \
                const T fyi = (T)floor((yr - y0) / (-pix));                                       \
                if (!(fxi >= (T)0 && fxi + (T)1 < (T)nx && fyi >= (T)0 && fyi + (T)1 < (T)nx))    \
                    continue;                                                                     \
                const int64_t xi = (int64_t)fxi, yi = (int64_t)fyi;                               \
                const T xw = (xr - ((T)xi * pix + x0)) / pix;                                     \
                const T yw = (yr - ((T)yi * (-pix) + y0)) / (-pix);                               \
                T val;                                                                            \
                if (cone) {                                                                       \
                    const T fzi = (T)floor((z - z0) / vh);                                        \
                    if (!(fzi >= (T)0 && fzi + (T)1 < (T)nz)) continue;                           \
                    const int64_t zi = (int64_t)fzi;                                              \
                    const T zw = (z - ((T)zi * vh + z0)) / vh;                                    \
                    const T *f = vol + (zi * nx + yi) * nx + xi;                                  \
                    const T lo = ((T)1 - xw) * ((T)1 - yw) * f[0] + xw * ((T)1 - yw) * f[1] +     \
                                 ((T)1 - xw) * yw * f[nx] + xw * yw * f[nx + 1];                  \
                    const T *g = f + plane;                                                       \
                    const T hi = ((T)1 - xw) * ((T)1 - yw) * g[0] + xw * ((T)1 - yw) * g[1] +     \
                                 ((T)1 - xw) * yw * g[nx] + xw * yw * g[nx + 1];                  \
                    val = (((T)1 - zw) * lo + zw * hi) * step * diag;                             \
                } else {                                                                          \
                    const T *f = vol + ((int64_t)r * nx + yi) * nx + xi;                          \
                    val = (((T)1 - xw) * ((T)1 - yw) * f[0] + xw * ((T)1 - yw) * f[1] +           \
                           ((T)1 - xw) * yw * f[nx] + xw * yw * f[nx + 1]) *                      \
                          step * diag;                                                            \
                }                                                                                 \
                acc = acc + val;                                                                  \
            }                                                                                     \
            out[ray] = acc;                                                                       \
        }                                                                                         \
        return 0;                                                                                 \
    }

DEFINE_FP(fp_oracle_f32, float)
DEFINE_FP(fp_oracle_f64, double)

/* ----------------------------------------------------------------------
 * IR back-projection.  sgm: [nrows][nviews][nu] (already multiplied by the
 * ray map); cosv/sinv: cos/sin of (angle - rotation) per view; cr/sr:
 * cos/sin of the rotation; vol: [nz][nx][nx], overwritten.
 * -------------------------------------------------------------------- */
#define DEFINE_BPIR(NAME, T)                                                                      \
    int NAME(const T *sgm, int nrows, int nviews, int nu, const T *cosv, const T *sinv, T cr,     \
             T sr, const T *u, const T *v, int nx, int nz, T pix, T vh, T cx, T cy, T cz,         \
             const T *center_z, T sid, T sdd, T det_h, int cone, T *vol) {                        \
        const T half = (T)((nx - 1) / 2.0), halfz = (T)((nz - 1) / 2.0);                          \
        const T u0 = u[0], du = u[1] - u[0];                                                      \
        const T v0 = v[0];                                                                        \
        const T dv = nrows > 1 ? v[1] - v[0] : (T)1;                                              \
        const T adv = dv < 0 ? -dv : dv;                                                          \
        const int64_t plane = (int64_t)nx * nx;                                                   \
        _Pragma("omp parallel for schedule(dynamic, 64)")                                         \
        for (int64_t col = 0; col < plane; ++col) {                                               \
            const int ix = (int)(col % nx), iy = (int)(col / nx);                                 \
            const T xa = pix * ((T)ix - half) + cx;                                               \
            const T ya = (-pix) * ((T)iy - half) + cy;                                            \
            const T x = xa * cr + ya * sr;                                                        \
            const T y = (-xa) * sr + ya * cr;                                                     \
            for (int iz = 0; iz < nz; ++iz) {                                                     \
                const T zb = ((T)iz - halfz) * vh;                                                \
                T acc = (T)0;                                                                     \
                int flag = 1;                                                                     \
                for (int j = 0; j < nviews; ++j) {                                                \
                    const T z = zb + (center_z ? center_z[j] : cz);                               \
                    const T L = sid - x * cosv[j] - y * sinv[j];                                  \
                    const T dxs = x - sid * cosv[j];                                              \
                    const T dys = y - sid * sinv[j];                                              \
                    const T adx = dxs < 0 ? -dxs : dxs, ady = dys < 0 ? -dys : dys;               \
                    const T den = adx > ady ? adx : ady;                                          \
                    const T df = (T)sqrt(dxs * dxs + dys * dys + z * z) / den;                    \
                    const T mag = sdd / L;                                                        \
                    const T uu = mag * ((-x) * sinv[j] + y * cosv[j]);                            \
                    const T uidx = (uu - u0) / du;                                                \
                    if (uidx < (T)0 || uidx + (T)1 > (T)(nu - 1)) {                               \
                        flag = 0;                                                                 \
                        continue;                                                                 \
                    }                                                                             \
                    const int uf = (int)floor(uidx);                                              \
                    const T ru = uidx - (T)uf;                                                    \
                    T val;                                                                        \
                    if (cone) {                                                                   \
                        const T vidx = (mag * z - v0) / det_h * adv / dv;                         \
                        const T vff = (T)floor(vidx);                                             \
                        if (vff < (T)0 || vff + (T)1 > (T)(nrows - 1)) {                          \
                            val = (T)0;                                                           \
                        } else {                                                                  \
                            const int64_t vf = (int64_t)vff;                                      \
                            const T rv = vidx - (T)vf;                                            \
                            const T *r0 = sgm + (vf * nviews + j) * nu;                           \
                            const T *r1 = r0 + (int64_t)nviews * nu;                              \
                            const T p0 = r0[uf] * ((T)1 - ru) + r0[uf + 1] * ru;                  \
                            const T p1 = r1[uf] * ((T)1 - ru) + r1[uf + 1] * ru;                  \
                            val = ((T)1 - rv) * p0 + rv * p1;                                     \
                        }                                                                         \
                    } else {                                                                      \
                        const T *r0 = sgm + ((int64_t)iz * nviews + j) * nu;                      \
                        val = ((T)1 - ru) * r0[uf] + ru * r0[uf + 1];                             \
                    }                                                                             \
                    acc = acc + val * df * pix;                                                   \
                }                                                                                 \
                vol[(int64_t)iz * plane + col] = flag ? acc : (T)0;                               \
            }                                                                                     \
        }                                                                                         \
        return 0;                                                                                 \
    }

DEFINE_BPIR(bpir_oracle_f32, float)
DEFINE_BPIR(bpir_oracle_f64, double)
