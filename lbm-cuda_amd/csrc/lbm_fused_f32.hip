// Instantiation of the k-step kernel for storage format FMT_F32 (a translation unit of its own so the
// formats compile in parallel).
#include "lbm_fused_impl.h"

namespace lbm {
template int launch_fused_fmt<FMT_F32>(const KPar &, int, int, int, int, int, hipStream_t, const uint8_t *,
                                       const void *, void *, const float *, int, float *, float *, float *);
} // namespace lbm
