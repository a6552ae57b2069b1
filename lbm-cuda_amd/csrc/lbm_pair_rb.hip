// Two-step tile kernel for the row-based 16-bit formats (deviations from a per-row base, lbm_params.row_base);
// kernels and launcher in lbm_pair_impl.h.
#include "lbm_pair_impl.h"

namespace lbm {

template int launch_pair_fmt<FMT_F16R>(const KPar &, int, int, int, hipStream_t, const uint8_t *, const void *, void *,
                                       const float *, float *, float *, float *);
template int launch_pair_fmt<FMT_BF16R>(const KPar &, int, int, int, hipStream_t, const uint8_t *, const void *,
                                        void *, const float *, float *, float *, float *);

} // namespace lbm
