// The four-step tile update: two even (A) and two odd (B) AA steps of a CW x CH tile in one launch,
// bit-identical to four single steps.  Each step has the single step's arithmetic; the tile only changes where
// the intermediate states live.
//
// Dependency radius.  An even step is local; an odd step gathers from the neighbours and scatters to them, so
// the state after it at a cell depends on the neighbours' previous states.  Four steps therefore need a 3-cell
// halo: A1 on every region cell (ring <= 3 around the core), B1 on ring <= 2, A2 on ring <= 1, B2 on the core.
//   A1  HBM -> registers -> collide -> reversed state into LDS (the two-step kernel's phase A on a wider ring)
//   B1  in LDS, in place: an odd AA step reads exactly the nine locations it writes (slot opp(q) at x - c_q is
//       slot q at x + c_q of the neighbour pair), so every cell of the ring updates concurrently without races
//   A2  in LDS, in place: local
//   B2  the core's odd step, gathered from LDS and scattered straight to the output; every output slot has one
//       writer (the AA ownership rule), so tiles need no ordering
// One barrier separates consecutive phases.  HBM sees the state read once (plus the halo rows neighbouring
// tiles share, mostly L2 hits) and written once per four steps.
//
// The reference's temporally blocked kernels (lbm_core_smemaa.cu:29-71,107-223,293-445;
// lbm_core_smemaa_endfused.cu:37-97,275-370) stage a T-deep halo in shared memory in the same way: A0 from HBM
// into registers, B-final scattered to HBM with single-owner writes.
#pragma once

#include <type_traits>

#include "lbm_pair_tile.h"

namespace lbm {
namespace quadk {

constexpr int R = 3; // halo rings

// Tile shapes by core rows CH: 128 x CH for 8 and 16 rows, 64 x CH for 12, 14, 20, 24 and 32; 1024-thread
// workgroups for 16 and 32 rows, 512 otherwise.  With FP32 populations 12 and 14 rows (46.6 and 51.8 KB of LDS)
// fit three workgroups per CU, 8, 20 and 24 rows two, 16 and 32 rows one.  Host code maps a run-time chunk with
// the same functions.
__host__ __device__ constexpr int quad_cw(int ch) { return ch == 8 || ch == 16 ? 128 : 64; }
__host__ __device__ constexpr int quad_nt(int ch) { return ch == 16 || ch == 32 ? 1024 : 512; }
__host__ __device__ constexpr bool quad_three(int ch) { return ch == 12 || ch == 14; }
template <int CH> struct Shape {
    static constexpr int CW = quad_cw(CH);
    static constexpr int NT = quad_nt(CH);
    static constexpr bool THREE = quad_three(CH); // three workgroups per CU: at most 85 VGPRs (6 waves per SIMD)
};

// The LDS image of a CW x CH tile's region, (CW + 6) x (CH + 6) cells.  The row stride is the region width, so
// the 64 consecutive cells a wave handles are 64 consecutive LDS words even when they span two rows.
template <int FMT, int CW, int CH, int NT> struct QuadLds {
    static constexpr int RW = CW + 2 * R, RH = CH + 2 * R;
    typename Conv<FMT>::T sf[9][RH][RW];
    uint8_t sm[RH][RW];                  // 1 = solid or outside the lattice
    alignas(8) uint8_t wsolid[NT / 64];  // per wave: 1 if any of its region cells is solid (no open boundary)
};

// INTERIOR: the region [x0 - 3, x0 + CW + 2] x [y0 - 3, y0 + CH + 2] lies in 1 <= x <= W - 3, 1 <= y <= H - 2, so
// every region cell is in the lattice, is updated by every step, is no inlet cell, feeds no outlet ghost, and
// no index wraps.
template <int CW, int CH> __device__ __forceinline__ bool quad_interior(int W, int H, int bx, int by)
{
    const int x0 = bx * CW, y0 = by * CH;
    return x0 >= 4 && x0 + CW + 5 <= W && y0 >= 4 && y0 + CH + 4 <= H;
}

template <int MODEL, bool PERIODIC, int FMT, int CW, int CH, int NT, bool INTERIOR>
__device__ __forceinline__ void quad_tile(const KPar &P, const uint8_t *__restrict__ mask, const void *__restrict__ fin,
                                          void *__restrict__ fout, const float *__restrict__ inlet_u,
                                          float *__restrict__ oux, float *__restrict__ ouy, float *__restrict__ orho,
                                          int bx, int by, QuadLds<FMT, CW, CH, NT> &L)
{
    using S = Store<FMT>;
    using C = Conv<FMT>;
    using T = typename C::T;
    constexpr unsigned B = S::BYTES;
    constexpr int RW = CW + 2 * R, RH = CH + 2 * R;
    constexpr int LS = RW;      // LDS row stride
    constexpr int PL = RH * LS; // LDS plane stride
    constexpr bool OPEN = !PERIODIC && !INTERIOR; // inlet, outlet and walls may lie in the region
    constexpr bool WRAP = PERIODIC && !INTERIOR;  // region indices may wrap around the lattice
    auto &sf = L.sf;
    auto &sm = L.sm;

    const int W = P.W, H = P.H;
    const unsigned plane = (unsigned)P.plane;
    const int x0 = bx * CW, y0 = by * CH;
    const rsrc_t rin = make_rsrc(fin, 9u * plane * B), rout = make_rsrc(fout, 9u * plane * B);
    const rsrc_t rm = make_rsrc(mask, plane);
    const int t = threadIdx.x;
    const CollideK K = opaque_consts(P); // kept in VGPRs across the unrolled cells
    bool any_solid = false;              // no open boundary: a region cell of this thread is solid

    // Cells the steps update (lattice coordinates of region cell (r, c)): wall rows, the outlet ghost column
    // and cells outside the lattice are not updated; solids are not either, the inlet column always is.
    auto gxy = [&](int r, int c, int &gx, int &gy) {
        gx = x0 - R + c;
        gy = y0 - R + r;
    };
    auto updated_rc = [&](int gx, int gy) { return (unsigned)(gy - 1) < (unsigned)(H - 2) && (unsigned)gx < (unsigned)(W - 1); };

    // ---- A1: even step of every region cell, reversed result into LDS (loads issued PD cells ahead)
    constexpr int NA = (RH * RW + NT - 1) / NT;
    constexpr int PD = NA < 4 ? NA : 4;
    float ga[NA][9];
    unsigned ma[NA], ia[NA];
    bool ina[NA], la[NA], ua[NA];
    auto load_cell = [&](int a) {
        const int e = t + a * NT;
        const int r = e / RW, c = e - r * RW;
        int gx = x0 - R + c, gy = y0 - R + r;
        bool in = e < RH * RW;
        if constexpr (WRAP) {
            if (W >= RW && H >= RH) { // the region lies at most one period outside the lattice
                gx = pairk::wrap1(gx, W);
                gy = pairk::wrap1(gy, H);
            } else {
                gx = pairk::wrap(gx, W);
                gy = pairk::wrap(gy, H);
            }
        } else if constexpr (OPEN) {
            in = in && (unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H;
        }
        const unsigned idx = in ? (unsigned)(gy * W + gx) : 0u;
        ma[a] = __builtin_amdgcn_raw_buffer_load_b8(rm, idx, 0, 0);
#pragma unroll
        for (int q = 0; q < 9; ++q) ga[a][q] = S::ld(rin, idx * B, (unsigned)q * plane * B, q);
        ina[a] = in;
        la[a] = OPEN && gx == 0;
        ua[a] = !OPEN || ((unsigned)(gy - 1) < (unsigned)(H - 2) && gx <= W - 2);
        ia[a] = (unsigned)gy;
    };
    auto collide_cell = [&](int a) {
        const int e = t + a * NT;
        const int r = e / RW, c = e - r * RW;
        const bool solid = !ina[a] || ma[a] != 0;
        if (e < RH * RW) sm[r][c] = solid ? 1 : 0;
        if constexpr (!OPEN) any_solid = any_solid || (ina[a] && ma[a] != 0);
        const bool inlet = la[a];
        if (!(ina[a] && ua[a] && (inlet || !solid))) return;
        float mac[3];
        collide_dev<MODEL>(P, ga[a], inlet, inlet ? inlet_u[ia[a]] : 0.0f, mac, -1.0f, K);
        T o[9];
        if constexpr (B == 2) {
#pragma unroll
            for (int q = 0; q < 8; q += 2) {
                const unsigned pk = C::st2(ga[a][q], q, ga[a][q + 1], q + 1, K);
                o[q] = (T)pk;
                o[q + 1] = (T)(pk >> 16);
            }
            o[8] = C::st(ga[a][8], 8, K);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) o[q] = C::st(ga[a][q], q, K);
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) sf[opp(q)][r][c] = o[q];
        if (OPEN && x0 - R + c == W - 2 && c + 1 < RW) { // reversed copy into the outlet ghost
#pragma unroll
            for (int q = 0; q < 9; ++q) sf[opp(q)][r][c + 1] = o[q];
        }
    };
#pragma unroll
    for (int a = 0; a < PD; ++a) load_cell(a);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        collide_cell(a);
        if (a + PD < NA) load_cell(a + PD);
    }
    if constexpr (!OPEN) {
        const bool wave_solid = __ballot(any_solid) != 0;
        L.wsolid[t / 64] = wave_solid ? 1 : 0;
    }
    __syncthreads();
    bool fluid_tile = false; // no open boundary and no solid cell in the region: no bounce-back anywhere
    if constexpr (!OPEN) {
        unsigned acc = 0;
#pragma unroll
        for (int w = 0; w < NT / 64; w += 8) {
            uint64_t v;
            __builtin_memcpy(&v, L.wsolid + w, sizeof(v));
            acc |= (unsigned)(v | (v >> 32));
        }
        fluid_tile = __builtin_amdgcn_readfirstlane(acc) == 0;
    }

    // Gather of an odd step for the cell whose 3 x 3 neighbourhood has its corner at LDS offset lb; returns
    // false when the cell is not updated.  solid[q]: the neighbour x + c_q is solid (never at the inlet's open
    // side).
    auto gather = [&](auto fluid_tag, int r, int c, int gx, int gy, float g[9], bool solid[9], bool &inlet) -> bool {
        constexpr bool FLUID = decltype(fluid_tag)::value;
        const int lb = (r - 1) * LS + (c - 1);
        const T *sfb = &sf[0][0][0] + lb;
        const uint8_t *smb = &sm[0][0] + lb;
        inlet = false;
        if constexpr (OPEN) {
            if (!updated_rc(gx, gy)) return false;
            inlet = gx == 0;
            if (!inlet && smb[LS + 1]) return false;
        } else if constexpr (!FLUID) {
            if (smb[LS + 1]) return false;
        }
#pragma unroll
        for (int q = 1; q < 9; ++q)
            solid[q] = !FLUID && smb[(1 + cy(q)) * LS + 1 + cx(q)] && !(OPEN && inlet && cx(q) < 0);
        solid[0] = false;
        g[0] = C::ld(sfb[LS + 1], 0);
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const int o = opp(q);
            const T own = sfb[q * PL + LS + 1], src = sfb[o * PL + (1 - cy(q)) * LS + (1 - cx(q))];
            g[q] = C::ld(solid[o] ? own : src, q);
        }
        return true;
    };

    // ---- B1: odd step of ring <= 2, in place in LDS
    auto phase_b1 = [&](auto fluid_tag) {
        constexpr int WB = CW + 4, NB = WB * (CH + 4);
        for (int e = t; e < NB; e += NT) {
            const int r = 1 + e / WB, c = 1 + (e - (e / WB) * WB);
            int gx, gy;
            gxy(r, c, gx, gy);
            float g[9], mac[3];
            bool solid[9], inlet;
            if (!gather(fluid_tag, r, c, gx, gy, g, solid, inlet)) continue;
            collide_dev<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac, -1.0f, K);
            T *const sfc = &sf[0][r][c];
            sfc[0] = C::st(g[0], 0, K);
#pragma unroll
            for (int q = 1; q < 9; ++q) {
                const int o = opp(q);
                const bool back = solid[q] || (OPEN && inlet && cx(q) < 0);
                // bounce-back: this cell's slot opp(q); otherwise the target's slot q
                const int off = back ? o * PL : q * PL + cy(q) * LS + cx(q);
                sfc[off] = C::st(g[q], q, K);
            }
            if (OPEN && gx == W - 2) {
                // the ghost (W-1, y) is a copy of this cell: stream its q = 3, 6, 7 into column W-2
                sfc[3 * PL] = C::st(g[3], 3, K);
                if (!solid[2]) sfc[6 * PL + LS] = C::st(g[6], 6, K);
                if (!solid[4]) sfc[7 * PL - LS] = C::st(g[7], 7, K);
            }
        }
    };
    if (fluid_tile)
        phase_b1(std::true_type{});
    else
        phase_b1(std::false_type{});
    __syncthreads();

    // ---- A2: even step of ring <= 1, in place in LDS (local)
    {
        constexpr int WA = CW + 2, NA2 = WA * (CH + 2);
        for (int e = t; e < NA2; e += NT) {
            const int r = 2 + e / WA, c = 2 + (e - (e / WA) * WA);
            int gx, gy;
            gxy(r, c, gx, gy);
            bool inlet = false;
            if constexpr (OPEN) {
                if (!updated_rc(gx, gy)) continue;
                inlet = gx == 0;
                if (!inlet && sm[r][c]) continue;
            } else {
                if (!fluid_tile && sm[r][c]) continue;
            }
            T *const sfc = &sf[0][r][c];
            float g[9], mac[3];
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] = C::ld(sfc[q * PL], q);
            collide_dev<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac, -1.0f, K);
            T o[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) o[q] = C::st(g[q], q, K);
#pragma unroll
            for (int q = 0; q < 9; ++q) sfc[opp(q) * PL] = o[q];
            if (OPEN && gx == W - 2) { // reversed copy into the outlet ghost
#pragma unroll
                for (int q = 0; q < 9; ++q) sfc[opp(q) * PL + 1] = o[q];
            }
        }
    }
    __syncthreads();

    // ---- B2: odd step of the core, scattered to the output
    auto phase_b2 = [&](auto fluid_tag) {
        constexpr bool FLUID = decltype(fluid_tag)::value;
        for (int e = t; e < CH * CW; e += NT) {
            const int r = R + e / CW, c = R + (e - (e / CW) * CW);
            int gx, gy;
            gxy(r, c, gx, gy);
            if constexpr (WRAP) {
                if (gx >= W || gy >= H) continue;
            }
            float g[9], mac[3];
            bool solid[9], inlet;
            if (!gather(fluid_tag, r, c, gx, gy, g, solid, inlet)) continue;
            collide_dev<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac, -1.0f, K);
            const unsigned idx = (unsigned)(gy * W + gx);
            const unsigned idxB = idx * B, pB = plane * B;
            if constexpr (FLUID && INTERIOR) {
                // every population streams to its target's slot q: one vector offset for all nine stores (the
                // lowest target's, x - 1, y - 1) and the wave-uniform rest in the scalar offset
                const unsigned lowB = (unsigned)(W + 1) * B;
                const unsigned vb = idxB - lowB;
                S::st(rout, vb, lowB, 0, g[0]);
#pragma unroll
                for (int q = 1; q < 9; ++q)
                    S::st(rout, vb, (unsigned)((cx(q) + cy(q) * W) * (int)B) + lowB + (unsigned)q * pB, q, g[q]);
            } else {
                const unsigned idx2 = idxB + 2u * pB;
                unsigned nbB[9];
#pragma unroll
                for (int q = 1; q < 9; ++q) {
                    if constexpr (WRAP)
                        nbB[q] = (unsigned)(pairk::wrap1(gy + cy(q), H) * W + pairk::wrap1(gx + cx(q), W)) * B;
                    else
                        nbB[q] = idxB + (unsigned)((cx(q) + cy(q) * W) * (int)B);
                }
                S::st(rout, idxB, 0, 0, g[0]);
#pragma unroll
                for (int q = 1; q < 9; ++q) {
                    const int o = opp(q);
                    const bool back = solid[q] || (OPEN && inlet && cx(q) < 0);
                    const unsigned off = q < o ? (back ? idx2 : nbB[q]) : (back ? idxB : nbB[q] + 2u * pB);
                    S::st(rout, off, (unsigned)(q < o ? q : o) * pB, q, g[q]);
                }
                if (OPEN && gx == W - 2) {
                    S::st(rout, idxB, 3u * pB, 3, g[3]);
                    if (!solid[2]) S::st(rout, nbB[2], 6u * pB, 6, g[6]);
                    if (!solid[4]) S::st(rout, nbB[4], 7u * pB, 7, g[7]);
                }
            }
            if (orho) {
                orho[idx] = mac[0];
                oux[idx] = mac[1];
                ouy[idx] = mac[2];
            }
        }
    };
    if (fluid_tile)
        phase_b2(std::true_type{});
    else
        phase_b2(std::false_type{});
}

template <int MODEL, bool PERIODIC, int FMT, int CW, int CH, int NT>
__device__ __forceinline__ void quad_tile_any(const KPar &P, const uint8_t *__restrict__ mask,
                                              const void *__restrict__ fin, void *__restrict__ fout,
                                              const float *__restrict__ inlet_u, float *__restrict__ oux,
                                              float *__restrict__ ouy, float *__restrict__ orho, int bx, int by,
                                              QuadLds<FMT, CW, CH, NT> &L)
{
    if (quad_interior<CW, CH>(P.W, P.H, bx, by))
        quad_tile<MODEL, PERIODIC, FMT, CW, CH, NT, true>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L);
    else
        quad_tile<MODEL, PERIODIC, FMT, CW, CH, NT, false>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L);
}

} // namespace quadk
} // namespace lbm
