// Per-sample sinogram weighting shared by the weighting kernel and the ramp
// filter's fused weighting pass.
#pragma once

#include "mct_internal.h"

namespace mct {

constexpr int kMaxBH = 16;

struct BHCoef {
    float c[kMaxBH];
};

// BHCorrection (run_mgfbp.py:946-954) then WeightSgm (run_mgfbp.py:956-1011)
// of one sample p at the detector element with positions (u, v) in scan view
// i, operation by operation in FP32; the library is built with
// -ffp-contract=off, so no product is fused into an FMA here.
__device__ __forceinline__ float weigh(float p, float u, float v, int i, const float *__restrict__ ang,
                                       float sdd, float scan, int short_scan, int curved,
                                       const BHCoef &bh, int bh_order) {
    if (bh_order > 0) {
        float acc = 0.f, pw = p;
        for (int t = 0; t < bh_order; ++t) {
            acc = acc + bh.c[t] * pw;
            pw = pw * p;
        }
        p = acc;
    }
    if (curved) {
        p = p * sdd * cosf((-1.f) * u / sdd) * sdd / sqrtf(sdd * sdd + v * v);
    } else {
        p = (p * sdd * sdd) / sqrtf(sdd * sdd + u * u + v * v);
    }
    if (short_scan) {
        float ascan = fabsf(scan);
        float num_rounds = floorf(ascan / (kPI * 2.f));
        float remain = ascan - num_rounds * kPI * 2.f;
        float beta = fabsf(ang[i] - ang[0]);
        float dir = ascan / scan;
        float gamma = curved ? ((-1.f) * u / sdd) * dir : atan2f((-1.f) * u, sdd) * dir;
        float gmax = remain - kPI;
        float w = 0.f;
        if (remain <= kPI && num_rounds == 0.f) {
            w = 1.f;
        } else if (remain <= kPI && num_rounds >= 1.f) {
            if (0.f <= beta && beta < remain)
                w = beta / remain;
            else if (remain <= beta && beta < 2.f * kPI * num_rounds)
                w = 1.f;
            else if (2.f * kPI * num_rounds <= beta && beta <= (2.f * kPI * num_rounds + remain))
                w = (2.f * kPI * num_rounds + remain - beta) / remain;
        } else if (kPI < remain && remain <= 2.f * kPI) {
            if (0.f <= beta && beta < (gmax - 2.f * gamma)) {
                w = sinf(kPI / 2.f * beta / (gmax - 2.f * gamma));
                w = w * w;
            } else if ((gmax - 2.f * gamma) <= beta && beta < (kPI * (2.f * num_rounds + 1.f) - 2.f * gamma)) {
                w = 1.f;
            } else if ((kPI * (2.f * num_rounds + 1.f) - 2.f * gamma) <= beta &&
                       beta <= (kPI * (2.f * num_rounds + 1.f) + gmax)) {
                w = sinf(kPI / 2.f * (kPI + gmax - (beta - kPI * 2.f * num_rounds)) / (gmax + 2.f * gamma));
                w = w * w;
            }
        } else {
            w = 1.f;
        }
        p *= w;
    }
    return p;
}

// What the ramp filter's input pass needs to weigh sample (row s, view i,
// element j) of a [n_v][nviews][n_u] block of views [view0, view0 + nviews).
struct WeighSpec {
    const float *u, *v, *ang;
    float sdd, scan;
    int short_scan, curved;
    BHCoef bh;
    int bh_order, nviews, view0;
};

}  // namespace mct
