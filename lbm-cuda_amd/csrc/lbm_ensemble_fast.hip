// Ensemble kernels with the fast arithmetic (lbm_params.fast = 1: the hardware's reciprocal and square root in the
// MRT collision): the MRT | MODEL_FAST instantiations of the persistent k = 1 kernel and the k = 2 and 4 tile
// kernels, in a translation unit of their own so that they compile in parallel with lbm_ensemble.hip.
#include "lbm_ensemble_impl.h"

namespace lbm {

constexpr int MODEL_MRT_FAST = MODEL_MRT | MODEL_FAST;

EnsKernel ens_kernel_fast(int fmt, int parity, bool periodic)
{
    if (fmt == LBM_FMT_F32) return ens_kernel_m<MODEL_MRT_FAST, FMT_F32>(parity, periodic);
    if (fmt == LBM_FMT_F16) return ens_kernel_m<MODEL_MRT_FAST, FMT_F16>(parity, periodic);
    return ens_kernel_m<MODEL_MRT_FAST, FMT_BF16>(parity, periodic);
}

PairKernel pair_kernel_fast(int fmt, bool periodic, int chi)
{
    if (fmt == LBM_FMT_F32)
        return periodic ? pair_kernel_c<MODEL_MRT_FAST, FMT_F32, true>(chi) : pair_kernel_c<MODEL_MRT_FAST, FMT_F32, false>(chi);
    if (fmt == LBM_FMT_F16)
        return periodic ? pair_kernel_c<MODEL_MRT_FAST, FMT_F16, true>(chi) : pair_kernel_c<MODEL_MRT_FAST, FMT_F16, false>(chi);
    return periodic ? pair_kernel_c<MODEL_MRT_FAST, FMT_BF16, true>(chi) : pair_kernel_c<MODEL_MRT_FAST, FMT_BF16, false>(chi);
}

PairKernel quad_kernel_fast(bool periodic, int qi)
{
    return periodic ? quad_kernel_c<MODEL_MRT_FAST, true>(qi) : quad_kernel_c<MODEL_MRT_FAST, false>(qi);
}

} // namespace lbm
