// Kernels of the batched ensembles (lbm_ensemble.hip; the compensated-sum instantiations live in
// lbm_ensemble_sums.hip so that they compile in parallel): the member descriptor, the persistent k = 1, 2 and 4
// kernels over a tile list spanning members, and templated selectors of their instantiations.
#pragma once

#include "lbm_aa_cell.h"
#include "lbm_common.h"
#include "lbm_pair_tile.h"
#include "lbm_quad_tile.h"

namespace lbm {

constexpr int ENS_NCH = 4;
constexpr int ENS_CH[ENS_NCH] = {8, 12, 16, 24}; // tile rows of the two-step list
constexpr int ENS_NQ = 4;
constexpr int ENS_QCH[ENS_NQ] = {8, 12, 20, 24}; // tile rows of the four-step lists (128 x 8, 64 x 12, 64 x 20, 64 x 24)

struct EnsDesc {
    KPar P;
    const uint8_t *mask;
    void *f[2]; // the state lives in f[side]; f[1] is the second buffer of the two-step path (may be NULL)
    const float *inlet_u;
    int tiles_x;
    int tile0;              // first 128 x 2 tile of this member
    int ptiles_x;           // 128-column tiles of the two-step list
    int ptile0[ENS_NCH];    // first 128 x CH tile of this member, per CH
    int qtiles_x[ENS_NQ];   // columns of tiles of the four-step lists
    int qtile0[ENS_NQ];     // first four-step tile of this member, per shape
};


// the member of list tile t among [m0, m1): uniform binary search over the members' first tiles
template <typename First>
__device__ __forceinline__ int ens_member(const EnsDesc *__restrict__ d, int m0, int m1, int t, First first)
{
    int lo = m0, hi = m1 - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first(d[mid]) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One single step of members [m0, m1) whose tiles are [t0, t1) of the k = 1 list, in their buffer `side`.
template <int MODEL, int FMT, int PARITY, bool PERIODIC>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_ensemble(const EnsDesc *__restrict__ d, int m0, int m1, int t0,
                                                              int t1, int side)
{
    const int G = gridDim.x, b = blockIdx.x;
    const int gper = G / 8;
    const int xcd = b % 8, l = b / 8;
    const int per = (t1 - t0 + 7) / 8;
    const int start = t0 + xcd * per, end = min(start + per, t1);
    for (int t = start + l; t < end; t += gper) {
        const EnsDesc &e = d[ens_member(d, m0, m1, t, [](const EnsDesc &x) { return x.tile0; })];
        const int lt = t - e.tile0;
        const int ty = lt / e.tiles_x, tx = lt - ty * e.tiles_x;
        aa_tile_cell<MODEL, FMT, PARITY, PERIODIC, false>(e.P, e.mask, e.f[side], e.inlet_u, nullptr, nullptr, nullptr,
                                                          tx, ty);
    }
}

// Tile of this workgroup in a grid of one workgroup per tile of an n-tile list: blocks are dealt round-robin to
// the 8 XCDs in dispatch order, so block b takes tile (b % 8) * (n / 8) + b / 8 and XCD i walks the i-th eighth of
// the list (consecutive tiles -- neighbouring rows of one member -- share an L2); the last n % 8 tiles keep
// dispatch order.
__device__ __forceinline__ int ens_tile(int n)
{
    const int b = (int)blockIdx.x, per = n / 8;
    return b < per * 8 ? (b % 8) * per + b / 8 : b;
}

// Two steps (even then odd) of members [m0, m1) whose tiles are [t0, t1) of the k = 2 list (tile rows
// ENS_CH[chi] = CH), from buffer `side` into buffer `side ^ 1`; one workgroup per tile (a persistent loop over
// tiles kept the loop's invariants live across the tile update: 88-135 VGPRs against the single lattice's 56-75).
template <int MODEL, int FMT, bool PERIODIC, int CH>
__global__ __launch_bounds__(pairk::NT) void k_ensemble_pair(const EnsDesc *__restrict__ d, int m0, int m1, int t0,
                                                             int t1, int side, int chi)
{
    __shared__ pairk::TileLds<FMT, CH> L;
    const int t = t0 + ens_tile(t1 - t0);
    const EnsDesc &e = d[ens_member(d, m0, m1, t, [chi](const EnsDesc &x) { return x.ptile0[chi]; })];
    const int lt = t - e.ptile0[chi];
    const int ty = lt / e.ptiles_x, tx = lt - ty * e.ptiles_x;
    pairk::pair_tile_any<MODEL, PERIODIC, FMT, CH>(e.P, e.mask, e.f[side], e.f[side ^ 1], e.inlet_u, nullptr, nullptr,
                                                   nullptr, tx, ty, L);
}

// Four steps of members [m0, m1) whose tiles are [t0, t1) of the four-step list qi (tile rows ENS_QCH[qi] = CH),
// from buffer `side` into buffer `side ^ 1`; one workgroup per tile.
#define LBM_ENS_QUAD_BODY                                                                                          \
    constexpr int CW = quadk::Shape<CH>::CW, NT = quadk::Shape<CH>::NT;                                           \
    __shared__ quadk::QuadLds<FMT_F32, CW, CH, NT> L;                                                             \
    const int t = t0 + ens_tile(t1 - t0);                                                                         \
    const EnsDesc &e = d[ens_member(d, m0, m1, t, [qi](const EnsDesc &x) { return x.qtile0[qi]; })];             \
    const int lt = t - e.qtile0[qi];                                                                              \
    const int ty = lt / e.qtiles_x[qi], tx = lt - ty * e.qtiles_x[qi];                                            \
    quadk::quad_tile_any<MODEL, PERIODIC, FMT_F32, CW, CH, NT>(e.P, e.mask, e.f[side], e.f[side ^ 1], e.inlet_u,   \
                                                                nullptr, nullptr, nullptr, tx, ty, L);

template <int MODEL, bool PERIODIC, int CH>
__global__ __launch_bounds__(quadk::Shape<CH>::NT) void k_ensemble_quad(const EnsDesc *__restrict__ d, int m0, int m1,
                                                                        int t0, int t1, int side, int qi)
{
    LBM_ENS_QUAD_BODY
}

// the shapes whose LDS fits three workgroups per CU: registers for six waves per SIMD
template <int MODEL, bool PERIODIC, int CH>
__global__ __launch_bounds__(quadk::Shape<CH>::NT) __attribute__((amdgpu_waves_per_eu(6))) void k_ensemble_quad3(
    const EnsDesc *__restrict__ d, int m0, int m1, int t0, int t1, int side, int qi)
{
    LBM_ENS_QUAD_BODY
}
#undef LBM_ENS_QUAD_BODY

using EnsKernel = void (*)(const EnsDesc *, int, int, int, int, int);
using PairKernel = void (*)(const EnsDesc *, int, int, int, int, int, int);

template <int MODEL, bool PERIODIC>
inline PairKernel quad_kernel_c(int qi)
{
    static_assert(ENS_QCH[1] == 12 && quadk::Shape<12>::THREE, "four-step list 1 is the three-workgroup shape");
    switch (qi) {
    case 0: return k_ensemble_quad<MODEL, PERIODIC, 8>;
    case 1: return k_ensemble_quad3<MODEL, PERIODIC, 12>;
    case 2: return k_ensemble_quad<MODEL, PERIODIC, 20>;
    default: return k_ensemble_quad<MODEL, PERIODIC, 24>;
    }
}

template <int MODEL, int FMT>
inline EnsKernel ens_kernel_m(int parity, bool periodic)
{
    if (parity)
        return periodic ? k_ensemble<MODEL, FMT, 1, true> : k_ensemble<MODEL, FMT, 1, false>;
    return periodic ? k_ensemble<MODEL, FMT, 0, true> : k_ensemble<MODEL, FMT, 0, false>;
}

template <int MODEL, int FMT, bool PERIODIC>
inline PairKernel pair_kernel_c(int chi)
{
    switch (chi) {
    case 0: return k_ensemble_pair<MODEL, FMT, PERIODIC, 8>;
    case 1: return k_ensemble_pair<MODEL, FMT, PERIODIC, 12>;
    case 2: return k_ensemble_pair<MODEL, FMT, PERIODIC, 16>;
    default: return k_ensemble_pair<MODEL, FMT, PERIODIC, 24>;
    }
}

// the compensated-sum (MODEL_MRT_SUMS) instantiations, defined in lbm_ensemble_sums.hip
EnsKernel ens_kernel_sums(int fmt, int parity, bool periodic);
PairKernel pair_kernel_sums(int fmt, bool periodic, int chi);
PairKernel quad_kernel_sums(bool periodic, int qi);
// the fast-arithmetic (MODEL_MRT | MODEL_FAST) instantiations, defined in lbm_ensemble_fast.hip
EnsKernel ens_kernel_fast(int fmt, int parity, bool periodic);
PairKernel pair_kernel_fast(int fmt, bool periodic, int chi);
PairKernel quad_kernel_fast(bool periodic, int qi);

} // namespace lbm
