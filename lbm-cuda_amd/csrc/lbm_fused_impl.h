// Temporally blocked k-step kernel (lbm_step_fused, lbm_step_fused_h16): k AA collide-stream steps per launch, bit-identical
// to k launches of the single-step kernel.
//
// The reference blocks in time with square CTA tiles and a T-deep halo in shared memory
// (lbm_core_smemaa.cu:29-71,107-109,293-295), which recomputes 1.47x (T=4) to 1.69x (T=6) of the useful
// cell updates.  Here a workgroup owns a strip of LW columns (core LW - 2k plus a k-column halo on each
// side) and sweeps down a chunk of rows; only the x halo and a k-row warm-up at each chunk end are
// recomputed.
//
// LDS holds one ring of R rows x LW columns x 9 populations, updated in place with the AA pattern, each
// row at its own time level.  Sweep iteration i loads row i (level 0), applies step s to row i - d_s with
// d_s = d_0 + 2s (d_0 = 1 + start parity), and writes row i - d_out (level k) to HBM.  With that skew
// every operation of one iteration touches its own rows: an even (A) step only its row, an odd (B) step
// rows r-1..r+1 of which it owns the slots it reads and writes (same ownership as the single-step kernel).
// One barrier per iteration orders the levels.  The next row's HBM loads are issued one iteration ahead
// into registers.  LDS holds the storage representation (FP32 f, or 16-bit deviations), so every
// step rounds exactly as the single-step kernel's HBM round trip does.
#pragma once

#include "lbm_common.h"

namespace lbm {

namespace fused {

constexpr int MAXK = FUSED_MAXK;
constexpr int PF = 4; // rows of HBM loads kept in flight ahead of the sweep (registers)

enum : uint8_t { F_SOLID = 1, F_COMPUTE = 2, F_INLET = 4, F_OUTADJ = 8 };

template <int K, int P0, int LW> struct Sched {
    static constexpr int D0 = 1 + P0;                   // lag of step 0 behind the load
    static constexpr int LASTB = (P0 + K - 1) & 1;      // last step is an odd (B) step
    static constexpr int DLAST = D0 + 2 * (K - 1);
    static constexpr int DOUT = DLAST + 1 + LASTB;      // lag of the HBM write-back
    static constexpr int R = DOUT + 1;                  // ring rows
    static constexpr int NT = (K * LW > 256) ? K * LW : 256;
    static constexpr int NLOAD = (9 * LW + NT - 1) / NT; // loaded floats per thread per row
    __host__ __device__ static constexpr int d(int s) { return D0 + 2 * s; }
    // column margin of step s: an odd step needs one valid column beyond its cells, and leaves a result
    // that is complete two columns further in (each odd step eats two halo columns, even steps none)
    __host__ __device__ static constexpr int nodd_before(int s) { return ((P0 + s) >> 1) - (P0 >> 1); }
    __host__ __device__ static constexpr int margin(int s) { return 2 * nodd_before(s) + ((P0 + s) & 1); }
    // halo (columns on each side, rows at each chunk end) consumed by the k steps
    static constexpr int HALO = 2 * nodd_before(K);
    static constexpr int CW = LW - 2 * HALO; // core columns written per strip
};
static_assert(Sched<8, 0, 128>::HALO == 8 && Sched<8, 1, 128>::HALO == 8 && Sched<1, 0, 128>::HALO == 0 &&
                  Sched<1, 1, 128>::HALO == 2,
              "halo");
static_assert(Sched<8, 0, 128>::margin(7) == 7 && Sched<8, 1, 128>::margin(7) == 8 &&
                  Sched<8, 1, 128>::margin(0) == 1,
              "margin");
static_assert(Sched<8, 0, 128>::R == 18 && Sched<8, 1, 128>::R == 18, "ring");

} // namespace fused

using namespace fused;

template <int MODEL, bool PERIODIC, int FMT, int K, int P0, int LW>
__global__ __launch_bounds__((Sched<K, P0, LW>::NT)) void k_fused(KPar P, const uint8_t *__restrict__ mask,
                                                                  const typename Conv<FMT>::T *__restrict__ fin,
                                                                  typename Conv<FMT>::T *__restrict__ fout,
                                                                  const float *__restrict__ inlet_u, int CH, int macros,
                                                                  float *__restrict__ oux, float *__restrict__ ouy,
                                                                  float *__restrict__ orho)
{
    using S = Sched<K, P0, LW>;
    using C = Conv<FMT>;
    using T = typename C::T;
    constexpr int R = S::R;
    constexpr int NT = S::NT;
    constexpr int CW = S::CW;
    constexpr int HALO = S::HALO;
    constexpr bool FINAL_REVERSED = ((P0 + K) & 1) != 0;
    __shared__ T sf[9][R][LW];
    __shared__ uint8_t sflag[R][LW];

    const int W = P.W, H = P.H, plane = P.plane;
    const int t = threadIdx.x;
    int bx, by;
    xcd_tile(bx, by);
    const int xs = bx * CW - HALO; // global column of strip column 0
    const int y0 = by * CH;        // first output row
    const int ch = min(CH, H - y0);
    const int ybeg = y0 - HALO;              // first row needed at level 0
    const int yend = y0 + ch + HALO - 1;     // last row needed at level 0
    const int niter = ch + HALO + S::DOUT;

    // ---- per-thread load and write-back slots: fixed for the whole sweep, so the divisions and the
    //      periodic column wrap are done once here rather than per row -------------------------------
    int lq[S::NLOAD], lc[S::NLOAD], lx[S::NLOAD]; // plane, strip column, global column (-1: outside)
#pragma unroll
    for (int n = 0; n < S::NLOAD; ++n) {
        const int e = t + n * NT;
        lq[n] = e / LW;
        lc[n] = e - (e / LW) * LW;
        int x = xs + lc[n];
        if (PERIODIC) x = ((x % W) + W) % W;
        lx[n] = (e < 9 * LW && x >= 0 && x < W) ? x : -1;
    }
    int fx = xs + t; // column whose flags this thread builds (t < LW)
    if (PERIODIC) fx = ((fx % W) + W) % W;
    constexpr int NOUT = (9 * CW + NT - 1) / NT;
    int oq[NOUT], oc[NOUT];
#pragma unroll
    for (int n = 0; n < NOUT; ++n) {
        const int e = t + n * NT;
        oq[n] = e < 9 * CW ? e / CW : -1;
        oc[n] = e - (e / CW) * CW;
    }

    // ---- level-0 rows travel HBM -> registers (PF rows ahead) -> LDS ring ------------------------
    T pre[PF][S::NLOAD];
    uint8_t pfl[PF];
    auto issue_load = [&](int j, T(&buf)[S::NLOAD], uint8_t &preflag) { // global row ybeg + j
        const int y = ybeg + j;
        const bool rowin = PERIODIC ? true : (y >= 0 && y < H);
        const int gy = PERIODIC ? ((y % H) + H) % H : y; // uniform
#pragma unroll
        for (int n = 0; n < S::NLOAD; ++n) {
            buf[n] = T(0);
            if (rowin && lx[n] >= 0) buf[n] = fin[lq[n] * plane + gy * W + lx[n]];
        }
        if (t < LW) {
            uint8_t fl = 0;
            if (PERIODIC) {
                fl = mask[gy * W + fx] ? F_SOLID : F_COMPUTE;
            } else if (rowin && fx >= 0 && fx < W) {
                const bool solid = mask[gy * W + fx] != 0;
                if (solid) fl |= F_SOLID;
                if (gy >= 1 && gy <= H - 2 && fx <= W - 2) {
                    if (fx == 0) fl |= F_COMPUTE | F_INLET;
                    else if (!solid) fl |= F_COMPUTE;
                    if (fx == W - 2) fl |= F_OUTADJ;
                }
            }
            preflag = fl;
        }
    };
    auto commit_load = [&](int j, const T(&buf)[S::NLOAD], uint8_t preflag) {
        const int slot = j % R;
#pragma unroll
        for (int n = 0; n < S::NLOAD; ++n)
            if (t + n * NT < 9 * LW) sf[lq[n]][slot][lc[n]] = buf[n];
        if (t < LW) sflag[slot][t] = preflag;
    };

    const int nload = yend - ybeg + 1;
#pragma unroll
    for (int u = 0; u < PF; ++u)
        if (u < nload) issue_load(u, pre[u], pfl[u]);
    // the u loop is unrolled so that every prefetch buffer is addressed statically (row i lives in
    // pre[i % PF]); the barrier at the end of each iteration waits for LDS only, so up to PF rows of HBM
    // loads stay in flight across iterations
    for (int i0 = 0; i0 < niter; i0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int i = i0 + u;
        if (i >= niter) break;
        // (1) row i enters the ring; (2) row i + PF is requested from HBM
        if (i < nload) commit_load(i, pre[u], pfl[u]);
        if (i + PF < nload) issue_load(i + PF, pre[u], pfl[u]);

        // (3) one cell of one step per thread
        // s is the same for the 64 lanes of a wave (LW is a multiple of 64): keep the step's row, ring slots
        // and margins in scalar registers
        const int s = __builtin_amdgcn_readfirstlane(t / LW), c = t - (t / LW) * LW;
        if (s < K) {
            const int jrow = i - S::d(s);                   // ring index of the row this step works on
            const int m = S::margin(s);
            const int y = ybeg + jrow;
            const bool odd = ((P0 + s) & 1) != 0;
            if (jrow >= m && ybeg + jrow <= yend - m && c >= m && c < LW - m) {
                const int sl = jrow % R;
                const uint8_t fl = sflag[sl][c];
                if (fl & F_COMPUTE) {
                    const bool inlet = (fl & F_INLET) != 0;
                    int gy = y;
                    if (PERIODIC) gy = ((y % H) + H) % H;
                    float g[9], mac[3];
                    if (!odd) {
#pragma unroll
                        for (int q = 0; q < 9; ++q) g[q] = C::ld(sf[q][sl][c], q);
                        collide_dev<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac);
                        T o[9];
#pragma unroll
                        for (int q = 0; q < 9; ++q) o[q] = C::st(g[q], q);
#pragma unroll
                        for (int q = 0; q < 9; ++q) sf[opp(q)][sl][c] = o[q];
                        if (!PERIODIC && (fl & F_OUTADJ) && c + 1 < LW) {
#pragma unroll
                            for (int q = 0; q < 9; ++q) sf[opp(q)][sl][c + 1] = o[q];
                        }
                    } else {
                        const int su = (jrow + 1) % R, sd = (jrow - 1) % R; // rows y+1, y-1
                        int rs[9], cs[9];
                        bool solid[9];
#pragma unroll
                        for (int q = 0; q < 9; ++q) {
                            rs[q] = cy(q) > 0 ? su : (cy(q) < 0 ? sd : sl);
                            cs[q] = c + cx(q);
                            const bool skip = !PERIODIC && inlet && cx(q) < 0;
                            solid[q] = (q == 0 || skip) ? false : (sflag[rs[q]][cs[q]] & F_SOLID) != 0;
                        }
                        g[0] = C::ld(sf[0][sl][c], 0);
#pragma unroll
                        for (int q = 1; q < 9; ++q) {
                            const int o = opp(q);
                            if (!PERIODIC && inlet && cx(q) > 0)
                                g[q] = 0.0f - wq(q); // arrives from x = -1; rebuilt by Zou-He
                            else
                                g[q] = C::ld(solid[o] ? sf[q][sl][c] : sf[o][rs[o]][cs[o]], q);
                        }
                        collide_dev<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac);
                        sf[0][sl][c] = C::st(g[0], 0);
#pragma unroll
                        for (int q = 1; q < 9; ++q) {
                            if (!PERIODIC && inlet && cx(q) < 0) continue;
                            if (solid[q])
                                sf[opp(q)][sl][c] = C::st(g[q], q);
                            else
                                sf[q][rs[q]][cs[q]] = C::st(g[q], q);
                        }
                        if (!PERIODIC && (fl & F_OUTADJ)) {
                            sf[3][sl][c] = C::st(g[3], 3);
                            if (!solid[2]) sf[6][su][c] = C::st(g[6], 6);
                            if (!solid[4]) sf[7][sd][c] = C::st(g[7], 7);
                        }
                    }
                    if (macros && s == K - 1) {
                        const int x = xs + c;
                        if (c >= HALO && c < LW - HALO && x < W && y >= y0 && y < y0 + ch) {
                            const int idx = y * W + x;
                            orho[idx] = mac[0];
                            oux[idx] = mac[1];
                            ouy[idx] = mac[2];
                        }
                    }
                }
            }
        }

        // (4) write back the row that has completed all k steps
        const int jout = i - S::DOUT;
        if (jout >= HALO && jout < HALO + ch) {
            const int sl = jout % R;
            const int y = ybeg + jout;
#pragma unroll
            for (int n = 0; n < NOUT; ++n) {
                const int q = oq[n], cc = oc[n];
                if (q < 0) continue;
                const int x = xs + HALO + cc;
                if (x >= W) continue;
                // The outlet ghost W-1 is part of cell W-2's state (its post-collision copy after an even
                // step), so the workgroup that owns W-2 writes it: the owner of W-1 may not have kept
                // W-2 valid through the last step.  A canonical ghost is never read and only written if
                // present.
                if (!PERIODIC && x == W - 1) continue;
                const T v = sf[q][sl][HALO + cc];
                fout[q * plane + y * W + x] = v;
                if (!PERIODIC && x == W - 2) {
                    if (FINAL_REVERSED && y >= 1 && y <= H - 2)
                        fout[q * plane + y * W + x + 1] = v;
                    else if (HALO + cc + 1 < LW)
                        fout[q * plane + y * W + x + 1] = sf[q][sl][HALO + cc + 1];
                }
            }
        }
        __syncthreads();
    }
    }
}

template <int MODEL, bool PERIODIC, int FMT, int K, int P0, int LW>
static void launch_fused_t(const KPar &P, hipStream_t st, const uint8_t *mask, const void *fin, void *fout,
                           const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
    using S = Sched<K, P0, LW>;
    using T = typename Conv<FMT>::T;
    dim3 grid((P.W + S::CW - 1) / S::CW, (P.H + CH - 1) / CH);
    hipLaunchKernelGGL((k_fused<MODEL, PERIODIC, FMT, K, P0, LW>), grid, dim3(S::NT), 0, st, P, mask,
                       static_cast<const T *>(fin), static_cast<T *>(fout), inlet_u, CH, rho != nullptr, ux, uy, rho);
}

template <int MODEL, bool PERIODIC, int FMT, int K, int LW>
static void launch_fused_p(const KPar &P, int parity, hipStream_t st, const uint8_t *mask, const void *fin, void *fout,
                           const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
    if (parity)
        launch_fused_t<MODEL, PERIODIC, FMT, K, 1, LW>(P, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
    else
        launch_fused_t<MODEL, PERIODIC, FMT, K, 0, LW>(P, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
}

template <int MODEL, bool PERIODIC, int FMT, int K>
static void launch_fused_w(const KPar &P, int lw, int parity, hipStream_t st, const uint8_t *mask, const void *fin,
                           void *fout, const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
    // 16-bit formats keep only the 128-column strip (their LDS ring is half the size)
    if (FMT == FMT_F32 && lw == 64)
        launch_fused_p<MODEL, PERIODIC, FMT, K, 64>(P, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
    else
        launch_fused_p<MODEL, PERIODIC, FMT, K, 128>(P, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
}

template <int MODEL, bool PERIODIC, int FMT>
static int launch_fused_k(const KPar &P, int k, int lw, int parity, hipStream_t st, const uint8_t *mask,
                          const void *fin, void *fout, const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
#define LBM_FUSED_CASE(KK)                                                                                              \
    case KK: launch_fused_w<MODEL, PERIODIC, FMT, KK>(P, lw, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho); break;
    switch (k) {
        LBM_FUSED_CASE(1)
        LBM_FUSED_CASE(2)
        LBM_FUSED_CASE(3)
        LBM_FUSED_CASE(4)
        LBM_FUSED_CASE(5)
        LBM_FUSED_CASE(6)
        LBM_FUSED_CASE(7)
        LBM_FUSED_CASE(8)
    default: return fail(LBM_ERR_ARG, "lbm_step_fused: k=%d outside [1, %d]", k, MAXK);
    }
#undef LBM_FUSED_CASE
    return LBM_OK;
}

// Launch of the k-step kernel with compensated sums (MODEL_MRT_SUMS; instantiated in lbm_fused_sums.hip).
template <int FMT>
int launch_fused_sums_fmt(const KPar &P, int bc, int k, int strip, int parity, hipStream_t st, const uint8_t *mask,
                          const void *fin, void *fout, const float *inlet_u, int CH, float *ux, float *uy, float *rho)
#ifdef LBM_FUSED_SUMS_TU
{
    return bc == LBM_BC_PERIODIC
               ? launch_fused_k<MODEL_MRT_SUMS, true, FMT>(P, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux, uy,
                                                           rho)
               : launch_fused_k<MODEL_MRT_SUMS, false, FMT>(P, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux,
                                                            uy, rho);
}
#else
    ;
#endif

// Launch of the k-step kernel for one storage format (instantiated in lbm_fused_<fmt>.hip).
template <int FMT>
int launch_fused_fmt(const KPar &P, int model, int bc, int k, int strip, int parity, hipStream_t st,
                     const uint8_t *mask, const void *fin, void *fout, const float *inlet_u, int CH, float *ux,
                     float *uy, float *rho)
{
    const bool per = bc == LBM_BC_PERIODIC;
    if (model == LBM_BGK)
        return per ? launch_fused_k<MODEL_BGK, true, FMT>(P, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux, uy,
                                                          rho)
                   : launch_fused_k<MODEL_BGK, false, FMT>(P, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux,
                                                           uy, rho);
    if (P.sums) return launch_fused_sums_fmt<FMT>(P, bc, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
    return per ? launch_fused_k<MODEL_MRT, true, FMT>(P, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho)
               : launch_fused_k<MODEL_MRT, false, FMT>(P, k, strip, parity, st, mask, fin, fout, inlet_u, CH, ux, uy,
                                                       rho);
}

} // namespace lbm
