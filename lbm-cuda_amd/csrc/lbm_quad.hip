// Four-step tile kernel: two even and two odd AA steps per launch (k = 4 from an even state), bit-identical
// to four launches of the single-step kernel; lbm_step_fused with strip = 0, k = 4 and chunk = tile rows
// (8, 12, 14, 16, 20, 24 or 32).  The tile update is quad_tile (lbm_quad_tile.h).  It halves the HBM bytes per step of the
// two-step kernel for FP32 populations, whose out-of-place two-step launches are bound by HBM: 72 B per cell
// per launch, 18 B per cell-step.
//
// Tile shapes, chosen by the launch configuration's chunk (quadk::Shape): 8 and 16 give 128 x chunk cores, 12, 14,
// 20, 24 and 32 give 64 x chunk cores, whose squarer regions carry less halo (the 3-ring halo costs (CW + 6)(CH + 6)
// region cells per CW x CH core in the first step).  FP32 LDS per workgroup and workgroups per CU (160 KB of LDS):
//   128 x 8   512 threads   69.4 KB  2        64 x 12   512 threads   46.6 KB  3
//   128 x 16  1024         109.0 KB  1        64 x 14   512           51.8 KB  3
//                                             64 x 20   512           67.3 KB  2
//                                             64 x 24   512           77.7 KB  2
//                                             64 x 32   1024          98.5 KB  1
// Three workgroups per CU hide the barriers between the phases best: 64 x 12 runs 8 % faster than 64 x 24 on the
// 4096 x 1024 cylinder although its halo makes 5.26 cell updates per core cell against 4.72.
// The grid is renumbered per XCD (xcd_tile) so that tiles sharing halo rows run on the same L2.
#include "lbm_quad_tile.h"

namespace lbm {
namespace quadk {

#define LBM_QUAD_KERNEL_BODY                                                                                       \
    constexpr int CW = Shape<CH>::CW, NT = Shape<CH>::NT;                                                         \
    __shared__ QuadLds<FMT, CW, CH, NT> L;                                                                        \
    int bx, by;                                                                                                   \
    xcd_tile(bx, by);                                                                                             \
    quad_tile_any<MODEL, PERIODIC, FMT, CW, CH, NT>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L);

template <int MODEL, bool PERIODIC, int FMT, int CH>
__global__ __launch_bounds__(Shape<CH>::NT) void k_quad(KPar P, const uint8_t *__restrict__ mask,
                                                        const void *__restrict__ fin, void *__restrict__ fout,
                                                        const float *__restrict__ inlet_u, float *__restrict__ oux,
                                                        float *__restrict__ ouy, float *__restrict__ orho)
{
    LBM_QUAD_KERNEL_BODY
}

// the shapes whose LDS fits three workgroups per CU: registers for six waves per SIMD
template <int MODEL, bool PERIODIC, int FMT, int CH>
__global__ __launch_bounds__(Shape<CH>::NT) __attribute__((amdgpu_waves_per_eu(6))) void k_quad3(
    KPar P, const uint8_t *__restrict__ mask, const void *__restrict__ fin, void *__restrict__ fout,
    const float *__restrict__ inlet_u, float *__restrict__ oux, float *__restrict__ ouy, float *__restrict__ orho)
{
    LBM_QUAD_KERNEL_BODY
}
#undef LBM_QUAD_KERNEL_BODY

template <int MODEL, bool PERIODIC, int FMT>
static void launch_ch(const KPar &P, int CH, hipStream_t st, const uint8_t *mask, const void *fin, void *fout,
                      const float *inlet_u, float *ux, float *uy, float *rho)
{
    auto go = [&](auto ch_tag) {
        constexpr int C = decltype(ch_tag)::value;
        constexpr int CW = Shape<C>::CW;
        const dim3 grid((P.W + CW - 1) / CW, (P.H + C - 1) / C);
        if constexpr (Shape<C>::THREE)
            hipLaunchKernelGGL((k_quad3<MODEL, PERIODIC, FMT, C>), grid, dim3(Shape<C>::NT), 0, st, P, mask, fin, fout,
                               inlet_u, ux, uy, rho);
        else
            hipLaunchKernelGGL((k_quad<MODEL, PERIODIC, FMT, C>), grid, dim3(Shape<C>::NT), 0, st, P, mask, fin, fout,
                               inlet_u, ux, uy, rho);
    };
    switch (CH) {
    case 32: go(std::integral_constant<int, 32>{}); break;
    case 24: go(std::integral_constant<int, 24>{}); break;
    case 20: go(std::integral_constant<int, 20>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    case 14: go(std::integral_constant<int, 14>{}); break;
    case 12: go(std::integral_constant<int, 12>{}); break;
    default: go(std::integral_constant<int, 8>{}); break;
    }
}

} // namespace quadk

// Launches the four-step tile kernel (k = 4 from an even state; CH = 8, 12, 14, 16, 20, 24 or 32, the shape
// table above).
template <int FMT>
int launch_quad_fmt(const KPar &P, int model, int bc, int CH, hipStream_t st, const uint8_t *mask, const void *fin,
                    void *fout, const float *inlet_u, float *ux, float *uy, float *rho)
{
    using namespace quadk;
    const bool per = bc == LBM_BC_PERIODIC;
    if (model == LBM_BGK) {
        if (per)
            launch_ch<MODEL_BGK, true, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_BGK, false, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
    } else if (P.sums) {
        if (per)
            launch_ch<MODEL_MRT_SUMS, true, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_MRT_SUMS, false, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
    } else if (P.fast) { // hardware reciprocal and square root (lbm_params.fast)
        if (P.clamp) {
            if (per)
                launch_ch<MODEL_MRT_CLAMP1 | MODEL_FAST, true, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
            else
                launch_ch<MODEL_MRT_CLAMP1 | MODEL_FAST, false, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
        } else {
            if (per)
                launch_ch<MODEL_MRT_CLAMP0 | MODEL_FAST, true, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
            else
                launch_ch<MODEL_MRT_CLAMP0 | MODEL_FAST, false, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
        }
    } else if (P.clamp) {
        if (per)
            launch_ch<MODEL_MRT_CLAMP1, true, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_MRT_CLAMP1, false, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
    } else {
        if (per)
            launch_ch<MODEL_MRT_CLAMP0, true, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_MRT_CLAMP0, false, FMT>(P, CH, st, mask, fin, fout, inlet_u, ux, uy, rho);
    }
    return LBM_OK;
}

template int launch_quad_fmt<FMT_F32>(const KPar &, int, int, int, hipStream_t, const uint8_t *, const void *, void *,
                                      const float *, float *, float *, float *);

} // namespace lbm
