// Exhaustive checks of the short arithmetic sequences the collision relies on, run over all 2^32 FP32 bit
// patterns on the device: recip_rn(x) against the general correctly rounded division __fdiv_rn(1, x), and
// sqrt_rn_normal(x) against the correctly rounded sqrtf for every normal x >= 1e-8 (its domain).  No
// reference counterpart: the reference compiles with -use_fast_math (lbm_smemaa_main.py:84) and rounds
// differently.
#include "lbm_common.h"

namespace lbm {

constexpr int SELFTEST_THREADS = 1 << 16; // each covers 2^16 consecutive bit patterns

__device__ __forceinline__ bool same_bits(float a, float b)
{
    return __builtin_bit_cast(uint32_t, a) == __builtin_bit_cast(uint32_t, b) || (a != a && b != b);
}

template <int WHAT>
__global__ __launch_bounds__(256) void k_selftest(uint32_t *__restrict__ bad, uint32_t *__restrict__ first)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t nbad = 0, f0 = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < (1u << 16); ++i) {
        const uint32_t bits = (t << 16) | i;
        const float x = __builtin_bit_cast(float, bits);
        bool ok = true;
        if constexpr (WHAT == 0) {
            ok = same_bits(recip_rn(x), __fdiv_rn(1.0f, x));
        } else {
            if (x >= 1e-8f && x <= 3.4028235e38f) ok = same_bits(sqrt_rn_normal(x), sqrtf(x));
        }
        if (!ok) {
            f0 = nbad == 0 ? bits : f0;
            ++nbad;
        }
    }
    bad[t] = nbad;
    first[t] = f0;
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_selftest(int what, uint32_t *bad, uint32_t *first, lbm_stream_t stream)
{
    if (what != 0 && what != 1) return fail(LBM_ERR_ARG, "lbm_selftest: what must be 0 (recip) or 1 (sqrt)");
    if (!bad || !first) return fail(LBM_ERR_ARG, "lbm_selftest: NULL output");
    const dim3 grid(SELFTEST_THREADS / 256);
    if (what == 0)
        hipLaunchKernelGGL(k_selftest<0>, grid, dim3(256), 0, to_hip(stream), bad, first);
    else
        hipLaunchKernelGGL(k_selftest<1>, grid, dim3(256), 0, to_hip(stream), bad, first);
    return check_launch("lbm_selftest");
}

int lbm_selftest_threads(void) { return SELFTEST_THREADS; }

} // extern "C"
