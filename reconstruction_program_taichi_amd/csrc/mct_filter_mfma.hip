// Ramp filter as a Toeplitz GEMM on MFMA (placeholder until the kernel lands).
#include "mct_internal.h"

namespace mct {
int filter_rows_mfma(mct_ctx *ctx, const float *in, float *out, int64_t n_rows, int n,
                     const float *taps, float scale, int method) {
    (void)ctx; (void)in; (void)out; (void)n_rows; (void)n; (void)taps; (void)scale;
    return set_error(MCT_ERR_UNSUPPORTED, "filter method %d is not available", method);
}
}  // namespace mct
