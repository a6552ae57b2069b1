// Ensemble kernels with compensated (Neumaier) sums of rho - 1 and j in the MRT collision (lbm_params.sums = 1):
// the MODEL_MRT_SUMS instantiations of the persistent k = 1, 2 and 4 kernels, in a translation unit of their own
// so that they compile in parallel with lbm_ensemble.hip.
#include "lbm_ensemble_impl.h"

namespace lbm {

EnsKernel ens_kernel_sums(int fmt, int parity, bool periodic)
{
    if (fmt == LBM_FMT_F32) return ens_kernel_m<MODEL_MRT_SUMS, FMT_F32>(parity, periodic);
    if (fmt == LBM_FMT_F16) return ens_kernel_m<MODEL_MRT_SUMS, FMT_F16>(parity, periodic);
    return ens_kernel_m<MODEL_MRT_SUMS, FMT_BF16>(parity, periodic);
}

PairKernel pair_kernel_sums(int fmt, bool periodic, int chi)
{
    if (fmt == LBM_FMT_F32)
        return periodic ? pair_kernel_c<MODEL_MRT_SUMS, FMT_F32, true>(chi) : pair_kernel_c<MODEL_MRT_SUMS, FMT_F32, false>(chi);
    if (fmt == LBM_FMT_F16)
        return periodic ? pair_kernel_c<MODEL_MRT_SUMS, FMT_F16, true>(chi) : pair_kernel_c<MODEL_MRT_SUMS, FMT_F16, false>(chi);
    return periodic ? pair_kernel_c<MODEL_MRT_SUMS, FMT_BF16, true>(chi) : pair_kernel_c<MODEL_MRT_SUMS, FMT_BF16, false>(chi);
}

PairKernel quad_kernel_sums(bool periodic, int qi)
{
    return periodic ? quad_kernel_c<MODEL_MRT_SUMS, true>(qi) : quad_kernel_c<MODEL_MRT_SUMS, false>(qi);
}

} // namespace lbm
