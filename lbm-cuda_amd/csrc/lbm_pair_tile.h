// The two-step tile update shared by the single-lattice kernel (lbm_pair_impl.h) and the batched ensemble
// (lbm_ensemble.hip): one even (A) and one odd (B) AA step of a 128 x CH tile, bit-identical to two single
// steps.  Phase A performs the even step of the tile plus a one-cell halo ring from HBM into LDS; phase B the
// odd step of the core, gathered from LDS and scattered straight to the output (every output slot has one
// writer).  The caller owns the LDS arrays and calls it from every thread of a 512-thread workgroup.
#pragma once

#include <type_traits>

#include "lbm_common.h"

namespace lbm {
namespace pairk {

constexpr int CW = 128;    // core columns per tile
constexpr int RW = CW + 2; // region columns (one halo column on each side)
constexpr int LS = 132;    // LDS row stride (elements)
constexpr int NT = 512;    // threads per workgroup
static_assert(NT / 64 == 8, "the per-wave votes are read as one 8-byte word");

// v mod n for v >= -n (one period at most below the lattice; region columns may lie several periods above
// it when the lattice is narrower than the region)
__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v % n : v); }
// v mod n for -n <= v < 2n (at most one period outside the lattice): no division
__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// The region rows' base values of a row-based 16-bit format (nothing for the other formats).  A staged row holds
// three groups of three directions by c_qy, each padded to 16 bytes: {0, 1, 3} (c_qy = 0), {2, 5, 6} (+1), {4, 7, 8}
// (-1).  The odd step of a cell on row r gathers group 0 of row r, group 1 of row r - 1 and group 2 of row r + 1,
// and scatters to group 0 of row r, group 1 of row r + 1 and group 2 of row r - 1: five 16-byte LDS reads.
template <bool RB, int RH> struct TileBaseRows {};
template <int RH> struct TileBaseRows<true, RH> {
    alignas(16) float base[RH][ROWB];
};
__host__ __device__ constexpr int bgrp(int q) { return cy(q) == 0 ? 0 : (cy(q) > 0 ? 1 : 2); }
__host__ __device__ constexpr int bidx(int q) { return q == 0 || q == 2 || q == 4 ? 0 : (q == 1 || q == 5 || q == 7 ? 1 : 2); }
// direction at position i < 3 of group g < 3 (the inverse of bgrp, bidx); -1 for the padding
__host__ __device__ constexpr int bdir(int g, int i)
{
    return i > 2 || g > 2 ? -1 : (g == 0 ? (i == 0 ? 0 : (i == 1 ? 1 : 3)) : (g == 1 ? (i == 0 ? 2 : (i == 1 ? 5 : 6))
                                                                                   : (i == 0 ? 4 : (i == 1 ? 7 : 8))));
}
static_assert(bdir(bgrp(7), bidx(7)) == 7 && bdir(bgrp(3), bidx(3)) == 3 && bdir(bgrp(6), bidx(6)) == 6, "groups");
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int FMT, int CH> struct TileLds : TileBaseRows<row_based<FMT>(), CH + 2> {
    typename Conv<FMT>::T sf[9][CH + 2][LS];
    uint8_t sm[CH + 2][LS];           // 1 = solid or outside the lattice
    alignas(8) uint8_t wsolid[NT / 64]; // per wave: 1 if any of its region cells is solid (no open boundary)
};

// INTERIOR: the tile's region (core plus halo ring) lies at least one cell inside the walls, the inlet column
// and the outlet cells x >= W - 2 (pair_interior), so every region cell is in the lattice, is updated by the
// even step, is no inlet cell and feeds no outlet ghost, and no index wraps.  That tile-uniform fact removes
// the per-cell boundary tests; the arithmetic is the same, so both forms give the same bits.
// Row-based 16-bit formats (row_based<FMT>): the region rows' base values are staged in LDS, read per lane by phase A
// (whose cells' rows vary within a wave) and with one address per wave by phase B (a wave covers half a core row); the
// LDS copy of the even step's result is stored relative to the base as in HBM, so both phases round as the single
// step does.  WIN (row-based formats, general tile form): the tile's region meets the base window (KPar::ww), whose
// cells take their base values from their own velocity instead of their row: the region's packed window velocities
// are staged in `wl` (RH x RW words of dynamic LDS, allocated by the launch when there is a window) with the row
// bases, and every base value a window cell needs is computed from them (vel_base).
template <int MODEL, bool PERIODIC, int FMT, int CH, bool INTERIOR, bool WIN = false>
__device__ __forceinline__ void pair_tile(const KPar &P, const uint8_t *__restrict__ mask, const void *__restrict__ fin,
                                          void *__restrict__ fout, const float *__restrict__ inlet_u,
                                          float *__restrict__ oux, float *__restrict__ ouy, float *__restrict__ orho,
                                          int bx, int by, TileLds<FMT, CH> &L, uint32_t *wl = nullptr)
{
    using S = Store<FMT>;
    using C = Conv<FMT>;
    using T = typename C::T;
    constexpr unsigned B = S::BYTES;
    constexpr bool RB = row_based<FMT>();
    constexpr int RH = CH + 2;
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
    const float one = RB ? opaque_one() : 1.0f;
    // row-based formats: thread t < 16 RH fetches position t % 16 of staged region row t / 16 (the grouped order of
    // TileBaseRows; padding and rows outside a channel lattice: 0, never used) before the population loads are
    // issued, and stores it to LDS after them
    float bstage = 0.0f;
    if constexpr (RB) {
        if (t < ROWB * RH) {
            const int r = t / ROWB, j = t % ROWB;
            const int q = bdir(j / 4, j % 4);
            int gy = y0 - 1 + r;
            bool ok = q >= 0;
            if constexpr (WRAP)
                gy = H >= RH ? wrap1(gy, H) : wrap(gy, H);
            else if constexpr (OPEN)
                ok = ok && (unsigned)gy < (unsigned)H;
            bstage = ok ? P.base[ROWB * gy + q] : 0.0f;
        }
    }
    // WIN: the packed velocities of the region cells (region coordinates as they are: the window lies inside the
    // lattice, so wrapped or outside cells are never window cells), fetched with the row bases
    constexpr int NWS = WIN ? (RH * RW + NT - 1) / NT : 1;
    uint32_t wstage[NWS];
    if constexpr (WIN) {
#pragma unroll
        for (int i = 0; i < NWS; ++i) {
            const int e = t + i * NT;
            const int r = e / RW, c = e - r * RW;
            wstage[i] = window_vel(P, x0 - 1 + c, y0 - 1 + r);
        }
    }
    // the staged group g (0, 1, 2) of region row rr, and the base value of direction q on region row rr
    auto bgroup = [&](int rr, int g) -> f32x4_t {
        if constexpr (RB)
            return reinterpret_cast<const f32x4_t *>(&L.base[rr][0])[g];
        else
            return f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    };

    // ---- phase A: even step of every region cell, reversed result into LDS.  A thread's HBM loads (mask byte
    //      + 9 populations per cell) are issued PD cells ahead of their use, and unconditionally, so their
    //      latencies overlap and no load waits behind a mask test; PD (4, or 2 for row-based formats) bounds the
    //      registers they hold.
    constexpr int NA = (RH * RW + NT - 1) / NT;
    // row-based: two cells in flight, for the registers of the base values held across the collision (measured
    // on the 4096 x 1024 cylinder: 1.8-2.2 % faster than three cells with the values read again after it)
    constexpr int PDMAX = RB ? 2 : 4;
    constexpr int PD = NA < PDMAX ? NA : PDMAX;
    float ga[NA][9];
    unsigned ma[NA], ia[NA];
    bool ina[NA], la[NA], ua[NA];
    auto load_cell = [&](int a) {
        const int e = t + a * NT;
        const int r = e / RW, c = e - r * RW;
        int gx = x0 - 1 + c, gy = y0 - 1 + r;
        bool in = e < RH * RW;
        if constexpr (WRAP) {
            // a region cell lies at most one period outside the lattice (x0 - 1 <= gx <= x0 + 128 < 2W for
            // W >= 130, likewise gy) unless the lattice is narrower or lower than the region (lattice-uniform)
            if (W >= RW && H >= RH) {
                gx = wrap1(gx, W);
                gy = wrap1(gy, H);
            } else {
                gx = wrap(gx, W);
                gy = wrap(gy, H);
            }
        } else if constexpr (OPEN) {
            // 0 <= v < n as one unsigned comparison (n > 0)
            in = in && (unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H;
        }
        // unconditional (cells outside the lattice read cell 0 and discard it), so no load waits in a branch
        const unsigned idx = in ? (unsigned)(gy * W + gx) : 0u;
        ma[a] = __builtin_amdgcn_raw_buffer_load_b8(rm, idx, 0, 0);
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            if constexpr (RB) // the raw bits; the base is added once the cell's row values are read from LDS
                ga[a][q] = __builtin_bit_cast(
                    float, (unsigned)__builtin_amdgcn_raw_buffer_load_b16(rin, idx * B, (unsigned)q * plane * B, 0));
            else
                ga[a][q] = S::ld(rin, idx * B, (unsigned)q * plane * B, q);
        }
        ina[a] = in;
        la[a] = OPEN && gx == 0;
        // rows and columns the even step updates (all of them when periodic): 1 <= gy <= H - 2, gx <= W - 2
        ua[a] = !OPEN || ((unsigned)(gy - 1) < (unsigned)(H - 2) && gx <= W - 2);
        ia[a] = (unsigned)gy; // the row, for the inlet speed
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
        // row-based formats: the base values of the cell's row, read from LDS before the collision and held across
        // it (the loads and the stores of an even step use the same nine values)
        float db[9] = {};
        auto read_base = [&]() {
            const f32x4_t b[3] = {bgroup(r, 0), bgroup(r, 1), bgroup(r, 2)};
#pragma unroll
            for (int q = 0; q < 9; ++q) db[q] = b[bgrp(q)][bidx(q)];
            if constexpr (WIN) { // region coordinates as they are: the window lies inside the lattice
                const bool inw = in_window(P, x0 - 1 + c, y0 - 1 + r);
                float ux, uy;
                unpack_vel(wl[r * RW + c], ux, uy);
                const float us = vel_usq15(ux, uy);
#pragma unroll
                for (int q = 0; q < 9; ++q) db[q] = inw ? vel_base(q, ux, uy, us) : db[q];
            }
        };
        if constexpr (RB) {
            read_base();
#pragma unroll
            for (int q = 0; q < 9; ++q) ga[a][q] = C::ldb((T)__builtin_bit_cast(unsigned, ga[a][q]), db[q], one);
        }
        collide_dev<MODEL>(P, ga[a], inlet, inlet ? inlet_u[ia[a]] : 0.0f, mac, -1.0f, K);
        T o[9];
        if constexpr (RB) {
            // the reversed slot opp(q) of row r holds direction q: base value db[q]
#pragma unroll
            for (int q = 0; q < 8; q += 2) {
                const unsigned pk = C::st2b(ga[a][q], q, db[q], ga[a][q + 1], q + 1, db[q + 1], K);
                o[q] = (T)pk;
                o[q + 1] = (T)(pk >> 16);
            }
            o[8] = C::stb(ga[a][8], 8, db[8], K);
        } else if constexpr (B == 2) {
            // pairs (0,1) (2,3) (4,5) (6,7) in one conversion each; the high halves go to LDS as they are
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
        if (OPEN && x0 - 1 + c == W - 2 && c + 1 < RW) { // reversed copy into the outlet ghost
#pragma unroll
            for (int q = 0; q < 9; ++q) sf[opp(q)][r][c + 1] = o[q];
        }
    };
#pragma unroll
    for (int a = 0; a < PD; ++a) load_cell(a);
    if constexpr (RB) {
        if (t < ROWB * RH) L.base[t / ROWB][t % ROWB] = bstage;
        if constexpr (WIN) {
#pragma unroll
            for (int i = 0; i < NWS; ++i)
                if (t + i * NT < RH * RW) wl[t + i * NT] = wstage[i];
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        collide_cell(a);
        if (a + PD < NA) load_cell(a + PD);
    }
    if constexpr (!OPEN) {
        // every lane of a wave writes the wave's vote (the same value) to the wave's byte
        const bool wave_solid = __ballot(any_solid) != 0;
        L.wsolid[t / 64] = wave_solid ? 1 : 0;
    }
    __syncthreads();

    // ---- phase B: odd step of the core cells, scattered to the output ------------------------------------
    // FLUID (interior or periodic tiles whose region holds no solid cell): no bounce-back anywhere, so the mask
    // tests and the gather / scatter selects drop out; the values moved and the arithmetic are the same.
    auto phase_b = [&](auto fluid_tag) {
        constexpr bool FLUID = decltype(fluid_tag)::value;
        for (int e = t; e < CH * CW; e += NT) {
            const int r = 1 + e / CW, c = 1 + (e - (e / CW) * CW);
            const int gx = x0 + c - 1, gy = y0 + r - 1;
            if constexpr (WRAP) {
                if (gx >= W || gy >= H) continue;
            } else if constexpr (OPEN) {
                // updated cells only: 1 <= gy <= H - 2 and gx <= W - 2 (walls, the outlet ghost and cells past the
                // lattice are not updated; gx, gy >= 0 here)
                if ((unsigned)(gy - 1) >= (unsigned)(H - 2) || gx >= W - 1) continue;
            }
            // LDS reads relative to the corner (r - 1, c - 1) of the cell's 3 x 3 neighbourhood: every access is
            // that base plus a non-negative constant, which the DS instructions take as an immediate offset
            const int lb = (r - 1) * LS + (c - 1);
            const T *sfb = &sf[0][0][0] + lb;
            const uint8_t *smb = &sm[0][0] + lb;
            auto at = [&](int k, int dy, int dx) -> T { return sfb[k * RH * LS + (1 + dy) * LS + (1 + dx)]; };
            bool inlet = false;
            if constexpr (OPEN) {
                inlet = gx == 0;
                if (!inlet && smb[LS + 1]) continue;
            } else if constexpr (!FLUID) {
                if (smb[LS + 1]) continue;
            }
            bool solid[9];
#pragma unroll
            for (int q = 1; q < 9; ++q)
                solid[q] = !FLUID && smb[(1 + cy(q)) * LS + 1 + cx(q)] && !(OPEN && inlet && cx(q) < 0);
            // row-based formats: base value of direction q on region row r + dy (the staged rows, with one address for
            // the whole wave).  The scatter's values -- direction q on row r + c_qy, groups 1 and 2 of rows r + 1 and
            // r - 1 -- are read here, before the gather and the collision, so that their latency overlaps them;
            // the others where used.  (Scalar loads of the three rows' values from the table instead ran the kernel
            // out of SGPRs: 35 spilled to VGPR lanes.)
            f32x4_t st_up = {}, st_dn = {};
            if constexpr (RB) {
                st_up = bgroup(r + 1, 1);
                st_dn = bgroup(r - 1, 2);
                asm volatile("" : "+v"(st_up), "+v"(st_dn)); // held in registers: not re-read after the collision
            }
            auto drow0 = [&](int dy, int q) -> float {
                if (dy != 0 && dy == cy(q)) // a scatter target's row: the values read above
                    return (dy > 0 ? st_up : st_dn)[bidx(q)];
                return bgroup(r + dy, bgrp(q))[bidx(q)];
            };
            // base value of direction q for a slot of cell (gx + dx, gy + dy) (WIN: that cell's own in the window)
            auto bv = [&](int q, int dx, int dy) -> float {
                if constexpr (WIN) {
                    float ux, uy;
                    unpack_vel(wl[(r + dy) * RW + (c + dx)], ux, uy);
                    return in_window(P, gx + dx, gy + dy) ? vel_base(q, ux, uy, vel_usq15(ux, uy)) : drow0(dy, q);
                } else {
                    return drow0(dy, q);
                }
            };
            float g[9], mac[3];
            if constexpr (RB)
                g[0] = C::ldb(at(0, 0, 0), bv(0, 0, 0), one);
            else
                g[0] = C::ld(at(0, 0, 0), 0);
#pragma unroll
            for (int q = 1; q < 9; ++q) {
                const int o = opp(q);
                // bounce-back: this cell's slot q; otherwise the source's slot opp(q) (both read, then selected).
                // At the inlet the q with c_qx > 0 come from x = -1 and are rebuilt by Zou-He, which never reads them.
                const T own = at(q, 0, 0), src = at(o, -cy(q), -cx(q));
                if constexpr (RB) // own: direction o on this row; source: direction q on row gy - c_qy
                    g[q] = C::ldb(solid[o] ? own : src, solid[o] ? bv(o, 0, 0) : bv(q, -cx(q), -cy(q)), one);
                else
                    g[q] = C::ld(solid[o] ? own : src, q);
            }
            collide_dev<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac, -1.0f, K);
            const unsigned idx = (unsigned)(gy * W + gx);
            const unsigned idxB = idx * B, pB = plane * B;
            if constexpr (FLUID && INTERIOR) {
                // every population streams to its target's slot q: one vector offset for all nine stores, the
                // lowest target's (x - 1, y - 1; interior cells have idx >= W + 1), and the wave-uniform rest,
                // (c_q,x + c_q,y W + W + 1) B + q plane B >= 0, as the scalar offset
                const unsigned lowB = (unsigned)(W + 1) * B;
                const unsigned vb = idxB - lowB;
                if constexpr (B == 2) {
                    // 16-bit: pairs (0,1) (2,3) (4,5) (6,7) converted in one instruction each; global stores take
                    // the high half as it is (the wave-uniform part of each address is the scalar base)
                    char *const ob = static_cast<char *>(fout);
                    auto put = [&](int q, uint16_t v) {
                        char *const bq = ob + ((cx(q) + cy(q) * W) * (int)B + (int)lowB) + (size_t)q * pB;
                        *reinterpret_cast<uint16_t *>(bq + vb) = v;
                    };
#pragma unroll
                    for (int q = 0; q < 8; q += 2) {
                        unsigned pk;
                        if constexpr (RB) // direction q lands on row gy + c_qy
                            pk = C::st2b(g[q], q, bv(q, cx(q), cy(q)), g[q + 1], q + 1,
                                         bv(q + 1, cx(q + 1), cy(q + 1)), K);
                        else
                            pk = C::st2(g[q], q, g[q + 1], q + 1, K);
                        put(q, (uint16_t)pk);
                        put(q + 1, (uint16_t)(pk >> 16));
                    }
                    if constexpr (RB)
                        put(8, C::stb(g[8], 8, bv(8, cx(8), cy(8)), K));
                    else
                        put(8, C::st(g[8], 8, K));
                } else {
                    S::st(rout, vb, lowB, 0, g[0]);
#pragma unroll
                    for (int q = 1; q < 9; ++q)
                        S::st(rout, vb, (unsigned)((cx(q) + cy(q) * W) * (int)B) + lowB + (unsigned)q * pB, q,
                              g[q]);
                }
            } else {
                const unsigned idx2 = idxB + 2u * pB;
                unsigned nbB[9];
#pragma unroll
                for (int q = 1; q < 9; ++q) {
                    if constexpr (WRAP)
                        // core cells lie in the lattice, so a neighbour is at most one period outside it
                        nbB[q] = (unsigned)(wrap1(gy + cy(q), H) * W + wrap1(gx + cx(q), W)) * B;
                    else
                        nbB[q] = idxB + (unsigned)((cx(q) + cy(q) * W) * (int)B);
                }
                // store of direction q to a slot whose row has the base values dr (row-based formats)
                auto put = [&](unsigned voff, unsigned soff, int q, float h, float dr) {
                    if constexpr (RB)
                        S::stb(rout, voff, soff, q, h, dr);
                    else
                        S::st(rout, voff, soff, q, h);
                };
                put(idxB, 0, 0, g[0], bv(0, 0, 0));
#pragma unroll
                for (int q = 1; q < 9; ++q) {
                    // bounce-back writes this cell's slot opp(q); populations leaving through the open inlet are parked
                    // in this cell's slot opp(q) (the canonical inlet slots 1, 5, 8, rebuilt by the next Zou-He).  The
                    // planes q and opp(q) differ by 2: min(q, opp(q)) * plane is the scalar offset, and the vector
                    // offset is this cell's or the target's offset, two planes up for the larger index.
                    const int o = opp(q);
                    const bool back = solid[q] || (OPEN && inlet && cx(q) < 0);
                    const unsigned off = q < o ? (back ? idx2 : nbB[q]) : (back ? idxB : nbB[q] + 2u * pB);
                    put(off, (unsigned)(q < o ? q : o) * pB, q, g[q], back ? bv(o, 0, 0) : bv(q, cx(q), cy(q)));
                }
                if (OPEN && gx == W - 2) {
                    // the ghost (W-1, y) is a copy of this cell: stream its q = 3, 6, 7 into column W-2
                    put(idxB, 3u * pB, 3, g[3], bv(3, 0, 0));
                    if (!solid[2]) put(nbB[2], 6u * pB, 6, g[6], bv(6, 0, 1));
                    if (!solid[4]) put(nbB[4], 7u * pB, 7, g[7], bv(7, 0, -1));
                }
            }
            if (orho) {
                orho[idx] = mac[0];
                oux[idx] = mac[1];
                ouy[idx] = mac[2];
            }
        }
    };
    if constexpr (!OPEN) { // interior tiles and periodic edge tiles: every region cell is in the lattice
        uint64_t votes;
        __builtin_memcpy(&votes, L.wsolid, sizeof(votes)); // one 8-byte LDS read
        if (__builtin_amdgcn_readfirstlane((unsigned)(votes | (votes >> 32))) == 0)
            phase_b(std::true_type{});
        else
            phase_b(std::false_type{});
    } else {
        phase_b(std::false_type{});
    }
}

// True when tile (bx, by) of a 128 x CH tiling is interior in the sense of pair_tile's INTERIOR: its region
// [x0 - 1, x0 + 128] x [y0 - 1, y0 + CH] lies in 1 <= x <= W - 3, 1 <= y <= H - 2.  (For a periodic lattice
// this is stricter than needed: it only has to avoid wrapping.)
template <int CH> __device__ __forceinline__ bool pair_interior(int W, int H, int bx, int by)
{
    const int x0 = bx * CW, y0 = by * CH;
    return x0 >= 2 && x0 + CW + 3 <= W && y0 >= 2 && y0 + CH + 2 <= H;
}

// Tile of this workgroup: the tiles meeting the base window (KPar::wt*, a rectangle of tiles) first, one per
// workgroup in dispatch order, so that their longer runs overlap the rest of the launch instead of trailing it;
// then the others in xcd_tile's banded order (eight contiguous bands, one per XCD), counted over the tiles outside
// the rectangle.  Without a window it is xcd_tile.
__device__ __forceinline__ void pair_tile_order(const KPar &P, int &bx, int &by)
{
    const int nbx = (int)gridDim.x, nb = (int)(gridDim.x * gridDim.y);
    const int b = (int)blockIdx.y * nbx + (int)blockIdx.x;
    const int nw = P.wtw * P.wth;
    if (b < nw) {
        bx = P.wtx0 + b % P.wtw;
        by = P.wty0 + b / P.wtw;
        return;
    }
    const int j = b - nw, per = (nb - nw) / 8;
    int t = j < per * 8 ? (j % 8) * per + j / 8 : j; // index among the tiles outside the rectangle, row-major
    const int above = P.wty0 * nbx, rest = nbx - P.wtw;
    if (t >= above) {
        const int k = t - above;
        if (k < P.wth * rest) { // a rectangle row (rest > 0 here)
            by = P.wty0 + k / rest;
            const int i = k % rest;
            bx = i < P.wtx0 ? i : i + P.wtw;
            return;
        }
        t += nw; // below the rectangle
    }
    bx = t % nbx;
    by = t / nbx;
}

// One tile, with the interior form where it applies (a workgroup-uniform branch).
template <int MODEL, bool PERIODIC, int FMT, int CH>
__device__ __forceinline__ void pair_tile_any(const KPar &P, const uint8_t *__restrict__ mask,
                                              const void *__restrict__ fin, void *__restrict__ fout,
                                              const float *__restrict__ inlet_u, float *__restrict__ oux,
                                              float *__restrict__ ouy, float *__restrict__ orho, int bx, int by,
                                              TileLds<FMT, CH> &L, uint32_t *wl = nullptr)
{
    if constexpr (row_based<FMT>()) {
        // tiles whose region meets the base window: the general form with per-cell base values
        const int x0 = bx * CW, y0 = by * CH;
        if (meets_window(P, x0 - 1, x0 + CW, y0 - 1, y0 + CH)) {
            pair_tile<MODEL, PERIODIC, FMT, CH, false, true>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L,
                                                             wl);
            return;
        }
    }
    if (pair_interior<CH>(P.W, P.H, bx, by))
        pair_tile<MODEL, PERIODIC, FMT, CH, true>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L);
    else
        pair_tile<MODEL, PERIODIC, FMT, CH, false>(P, mask, fin, fout, inlet_u, oux, ouy, orho, bx, by, L);
}

} // namespace pairk
} // namespace lbm
