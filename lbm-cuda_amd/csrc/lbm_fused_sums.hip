// Instantiation of the k-step strip kernel with compensated (Neumaier) sums of rho - 1 and j
// (lbm_params.sums = 1) for the three storage formats, in a translation unit of its own so that it compiles in
// parallel with the plain instantiations.
#define LBM_FUSED_SUMS_TU
#include "lbm_fused_impl.h"

namespace lbm {
template int launch_fused_sums_fmt<FMT_F32>(const KPar &, int, int, int, int, hipStream_t, const uint8_t *,
                                            const void *, void *, const float *, int, float *, float *, float *);
template int launch_fused_sums_fmt<FMT_F16>(const KPar &, int, int, int, int, hipStream_t, const uint8_t *,
                                            const void *, void *, const float *, int, float *, float *, float *);
template int launch_fused_sums_fmt<FMT_BF16>(const KPar &, int, int, int, int, hipStream_t, const uint8_t *,
                                             const void *, void *, const float *, int, float *, float *, float *);
} // namespace lbm
