// Two-step tile kernel for FP32 populations and the plain 16-bit formats (kernels and launcher in
// lbm_pair_impl.h).
#include "lbm_pair_impl.h"

namespace lbm {

template int launch_pair_fmt<FMT_F32>(const KPar &, int, int, int, hipStream_t, const uint8_t *, const void *, void *,
                                      const float *, float *, float *, float *);
template int launch_pair_fmt<FMT_F16>(const KPar &, int, int, int, hipStream_t, const uint8_t *, const void *, void *,
                                      const float *, float *, float *, float *);
template int launch_pair_fmt<FMT_BF16>(const KPar &, int, int, int, hipStream_t, const uint8_t *, const void *,
                                       void *, const float *, float *, float *, float *);

} // namespace lbm
