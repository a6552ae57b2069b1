// placeholder, replaced below
#include "lbm_common.h"
extern "C" {
int lbm_fused_max_k(void) { return 0; }
int lbm_step_fused(const lbm_params *, const uint8_t *, const float *, float *, const float *, int, int, float *, float *,
                   float *, lbm_stream_t)
{
    return lbm::fail(LBM_ERR_UNSUPPORTED, "lbm_step_fused: not built");
}
}
