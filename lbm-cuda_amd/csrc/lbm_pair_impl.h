// Two-step tile kernel: one even (A) and one odd (B) AA step in one launch, bit-identical to two launches of
// the single-step kernel, for lbm_step_fused / lbm_step_fused_h16 with k = 2 from an even state (selected by a
// launch configuration with strip = 0; chunk = tile rows, 8, 12, 16 or 24).
//
// The reference fuses T steps in square shared-memory tiles (lbm_core_smemaa.cu:29-71,107-223,293-445).  Here
// a workgroup owns a core of 128 x CH cells:
//   phase A  loads the core plus a one-cell halo ring (130 x (CH + 2) cells) straight from HBM, performs the
//            local even step of every one of them and keeps the reversed result in LDS (storage format, so
//            the rounding is the single-step kernel's); the outlet ghost receives its copy there too;
//   phase B  after one barrier, performs the odd step of the core cells: gather from LDS, collide, and scatter
//            directly to the output buffer.  Every output slot has exactly one writer (the AA ownership rule
//            of the single step), so no LDS staging or inter-tile ordering is needed for the output.
// Per two steps HBM sees (130 (CH + 2)) / (128 CH) of the state read and the state written once; the halo costs
// the same fraction of extra even-step collisions.
//
// The launcher template is instantiated per storage format in lbm_pair.hip (FP32, FP16, BF16) and lbm_pair_rb.hip
// (the row-based 16-bit formats), so the translation units compile in parallel.
#pragma once

#include "lbm_pair_tile.h"

namespace lbm {
namespace pairk {

template <int MODEL, bool PERIODIC, int FMT, int CH>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_pair(KPar P, const uint8_t *__restrict__ mask, const void *__restrict__ fin,
                                             void *__restrict__ fout, const float *__restrict__ inlet_u,
                                             float *__restrict__ oux, float *__restrict__ ouy,
                                             float *__restrict__ orho)
{
    __shared__ TileLds<FMT, CH> L;
    extern __shared__ uint32_t pair_wl[]; // with a base window: (CH + 2) x RW packed velocities (pair_tile WIN)
    int bx, by;
    if constexpr (row_based<FMT>())
        pair_tile_order(P, bx, by);
    else
        xcd_tile(bx, by);
    pair_tile_any<MODEL, PERIODIC, FMT, CH>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L, pair_wl);
}

template <int MODEL, bool PERIODIC, int FMT>
static void launch_ch(const KPar &P, int CH, dim3 grid, unsigned dyn, hipStream_t st, const uint8_t *mask,
                      const void *fin,
                      void *fout, const float *inlet_u, float *ux, float *uy, float *rho)
{
    switch (CH) {
    case 24:
        hipLaunchKernelGGL((k_pair<MODEL, PERIODIC, FMT, 24>), grid, dim3(NT), dyn, st, P, mask, fin, fout, inlet_u, ux,
                           uy, rho);
        break;
    case 16:
        hipLaunchKernelGGL((k_pair<MODEL, PERIODIC, FMT, 16>), grid, dim3(NT), dyn, st, P, mask, fin, fout, inlet_u, ux,
                           uy, rho);
        break;
    case 12:
        hipLaunchKernelGGL((k_pair<MODEL, PERIODIC, FMT, 12>), grid, dim3(NT), dyn, st, P, mask, fin, fout, inlet_u, ux,
                           uy, rho);
        break;
    default:
        hipLaunchKernelGGL((k_pair<MODEL, PERIODIC, FMT, 8>), grid, dim3(NT), dyn, st, P, mask, fin, fout, inlet_u, ux,
                           uy, rho);
    }
}

} // namespace pairk

// Launches the two-step tile kernel (k = 2 from an even state; CH = 8, 12, 16 or 24 core rows per tile).
template <int FMT>
int launch_pair_fmt(const KPar &P, int model, int bc, int CH, hipStream_t st, const uint8_t *mask, const void *fin,
                    void *fout, const float *inlet_u, float *ux, float *uy, float *rho)
{
    using namespace pairk;
    const dim3 grid((P.W + CW - 1) / CW, (P.H + CH - 1) / CH);
    const bool per = bc == LBM_BC_PERIODIC;
    // a base window: the rectangle of tiles whose region [x0 - 1, x0 + CW] x [y0 - 1, y0 + CH] meets it (the tiles
    // that take the window form, dispatched first) and the dynamic LDS for their staged velocities
    KPar Q = P;
    Q.wtx0 = Q.wty0 = Q.wtw = Q.wth = 0;
    unsigned dyn = 0;
    if (row_based<FMT>() && P.ww > 0) {
        const int bx0 = P.wx0 <= CW ? 0 : (P.wx0 - 1) / CW, by0 = P.wy0 <= CH ? 0 : (P.wy0 - 1) / CH;
        const int bx1 = std::min((int)grid.x - 1, (P.wx0 + P.ww) / CW);
        const int by1 = std::min((int)grid.y - 1, (P.wy0 + P.wh) / CH);
        Q.wtx0 = bx0;
        Q.wty0 = by0;
        Q.wtw = bx1 - bx0 + 1;
        Q.wth = by1 - by0 + 1;
        dyn = (unsigned)((CH + 2) * RW * sizeof(uint32_t));
    }
    if (model == LBM_BGK) {
        if (per)
            launch_ch<MODEL_BGK, true, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_BGK, false, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
    } else if (P.sums) { // compensated sums read the clamp setting at run time
        if (per)
            launch_ch<MODEL_MRT_SUMS, true, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_MRT_SUMS, false, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
    } else if (P.fast) { // hardware reciprocal and square root (lbm_params.fast)
        if (P.clamp) {
            if (per)
                launch_ch<MODEL_MRT_CLAMP1 | MODEL_FAST, true, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
            else
                launch_ch<MODEL_MRT_CLAMP1 | MODEL_FAST, false, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
        } else {
            if (per)
                launch_ch<MODEL_MRT_CLAMP0 | MODEL_FAST, true, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
            else
                launch_ch<MODEL_MRT_CLAMP0 | MODEL_FAST, false, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
        }
    } else if (P.clamp) { // the clamp setting as a template argument: no run-time select per cell
        if (per)
            launch_ch<MODEL_MRT_CLAMP1, true, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_MRT_CLAMP1, false, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
    } else {
        if (per)
            launch_ch<MODEL_MRT_CLAMP0, true, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            launch_ch<MODEL_MRT_CLAMP0, false, FMT>(Q, CH, grid, dyn, st, mask, fin, fout, inlet_u, ux, uy, rho);
    }
    return LBM_OK;
}

} // namespace lbm
