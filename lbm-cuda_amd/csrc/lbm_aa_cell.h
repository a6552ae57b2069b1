// One in-place AA step of a single cell, shared by the single-lattice kernel (lbm_step.hip) and the
// persistent ensemble kernel (lbm_ensemble.hip).
#pragma once

#include "lbm_common.h"

namespace lbm {

// One AA step of cell (x, y); cells outside the lattice return at once.  EXT: the stress moments relax at
// w_ext when w_ext >= 0 (a learned closure's rate), else by the model's own rule.  INTERIOR: the caller knows
// that 1 <= x <= W - 3 and 1 <= y <= H - 2 (no wall row, inlet, outlet ghost or ghost source, and no neighbour
// index wraps), which removes the per-cell boundary tests; the arithmetic is the same.  Row-based 16-bit formats
// (row_based<FMT>) add the base value of each slot's row and direction on load and subtract it on store: the even step
// reads and writes row y, the odd step gathers from rows y - c_qy (bounce-back: this cell's row) and scatters to rows
// y + c_qy; the three rows' values sit in scalar registers (a wave covers part of one row).  NEARW: some cell whose
// slots this cell reads or writes may lie in the base window (KPar::ww): each access then takes that cell's own base
// values where it does.
template <int MODEL, int FMT, int PARITY, bool PERIODIC, bool MACROS, bool EXT = false, bool INTERIOR = false,
          bool NEARW = false>
__device__ __forceinline__ void aa_cell(const KPar &P, const uint8_t *__restrict__ mask, void *__restrict__ f,
                                        const float *__restrict__ inlet_u, float *__restrict__ oux,
                                        float *__restrict__ ouy, float *__restrict__ orho, int x, int y,
                                        float w_ext = -1.0f)
{
    using S = Store<FMT>;
    constexpr unsigned B = S::BYTES;
    constexpr bool RB = row_based<FMT>();
    constexpr bool OPEN = !PERIODIC && !INTERIOR; // walls, inlet and outlet may be in reach
    constexpr bool WRAP = PERIODIC && !INTERIOR;  // neighbour indices may wrap around the lattice
    if (!INTERIOR && (x >= P.W || y >= P.H)) return;
    const int W = P.W;
    const unsigned plane = (unsigned)P.plane;
    const unsigned idx = (unsigned)(y * W + x);
    const rsrc_t rf = make_rsrc(f, 9u * plane * B);
    const rsrc_t rm = make_rsrc(mask, plane);
    auto solid_at = [&](unsigned cell) { return __builtin_amdgcn_raw_buffer_load_b8(rm, cell, 0, 0) != 0; };
    bool inlet = false;
    if constexpr (OPEN) {
        // wall rows and the outlet ghost are never updated (lbm_core4.cu:20-24)
        if (y == 0 || y == P.H - 1 || x == W - 1) return;
        inlet = x == 0;
        if (!inlet && solid_at(idx)) return;
    } else {
        if (solid_at(idx)) return;
    }
    float g[9], mac[3];
    unsigned nb[9], nbB[9], nb2[9];
    bool solid[9];
    // row-based formats: base values of rows y - 1, y, y + 1 (wrapped when periodic; channel cells updated here
    // have 1 <= y <= H - 2)
    float dm[9] = {}, d0[9] = {}, dp[9] = {};
    float one = 1.0f;
    if constexpr (RB) {
        one = opaque_one();
        base_row_s(P.base, y, d0);
        if constexpr (PARITY == 1) {
            base_row_s(P.base, WRAP && y == 0 ? P.H - 1 : y - 1, dm);
            base_row_s(P.base, WRAP && y == P.H - 1 ? 0 : y + 1, dp);
        }
    }
    // base value of direction q on the row y + dy (dy = -1, 0, 1)
    auto drow = [&](int dy, int q) { return dy < 0 ? dm[q] : (dy > 0 ? dp[q] : d0[q]); };
    // base value of direction q for a slot of cell (x + dx, y + dy)
    auto bv = [&](int q, int dx, int dy) -> float {
        if constexpr (NEARW)
            return window_or(P, x + dx, y + dy, q, drow(dy, q));
        else
            return drow(dy, q);
    };
    if constexpr (PARITY == 0) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            if constexpr (RB)
                g[q] = S::ldb(rf, idx * B, (unsigned)q * plane * B, bv(q, 0, 0), one);
            else
                g[q] = S::ld(rf, idx * B, (unsigned)q * plane * B, q);
    } else {
        nb[0] = idx;
        solid[0] = false;
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            if constexpr (WRAP) {
                int nx = x + cx(q), ny = y + cy(q);
                nx = nx < 0 ? nx + W : (nx >= W ? nx - W : nx);
                ny = ny < 0 ? ny + P.H : (ny >= P.H ? ny - P.H : ny);
                nb[q] = (unsigned)(ny * W + nx);
            } else {
                // x = -1 only at the inlet: the previous row's last cell, or past the end of the mask (reads
                // 0); either way the result is discarded below
                nb[q] = idx + (unsigned)(cx(q) + cy(q) * W);
            }
            // all eight mask loads are issued unconditionally, so that they share one wait
            const bool s_nb = solid_at(nb[q]);
            solid[q] = s_nb && !(OPEN && inlet && cx(q) < 0);
        }
        // Bounce-back reads this cell's slot q, otherwise the source's slot opp(q).  The choice is a per-lane
        // select of the vector offset: a per-lane scalar offset would force a loop over its distinct values
        // around every load.  The two candidates' planes q and opp(q) differ by 2, so the wave-uniform part
        // min(q, opp(q)) * plane goes into the scalar offset and the vector offsets need only this cell's
        // offset, the same two planes up (idx2), and the neighbours' offsets, two planes up for q = 3, 4, 7, 8
        // (nb2): 13 additions for the gather and the scatter together.  At the inlet the q with c_qx > 0
        // arrive from x = -1; the value read for them (from an in-range or past-the-end offset, both harmless)
        // is replaced by Zou-He, which never reads it.
        const unsigned idxB = idx * B, pB = plane * B;
        const unsigned idx2 = idxB + 2u * pB;
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            if constexpr (WRAP)
                nbB[q] = nb[q] * B;
            else
                nbB[q] = idxB + (unsigned)((cx(q) + cy(q) * W) * (int)B);
            nb2[q] = opp(q) < q ? nbB[q] + 2u * pB : 0u; // used for q = 3, 4, 7, 8 only
        }
        if constexpr (RB)
            g[0] = S::ldb(rf, idxB, 0, bv(0, 0, 0), one);
        else
            g[0] = S::ld(rf, idxB, 0, 0);
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const int o = opp(q);
            // o < q: slot o at the neighbour = nbB[o] + o pB, slot q here = idx2 + o pB
            // o > q: slot o at the neighbour = nb2[o] + q pB, slot q here = idxB + q pB
            const unsigned off = o < q ? (solid[o] ? idx2 : nbB[o]) : (solid[o] ? idxB : nb2[o]);
            if constexpr (RB) // bounce-back: this row's slot holding direction o; else the source row's, direction q
                g[q] = S::ldb(rf, off, (unsigned)(o < q ? o : q) * pB, solid[o] ? bv(o, 0, 0) : bv(q, -cx(q), -cy(q)),
                              one);
            else
                g[q] = S::ld(rf, off, (unsigned)(o < q ? o : q) * pB, q);
        }
    }
    collide_dev<MODEL, EXT>(P, g, inlet, inlet ? inlet_u[y] : 0.0f, mac, w_ext);
    // store of direction q to a slot whose row has the base values dr (row-based formats)
    auto put = [&](unsigned voff, unsigned soff, int q, float h, float dr) {
        if constexpr (RB)
            S::stb(rf, voff, soff, q, h, dr);
        else
            S::st(rf, voff, soff, q, h);
    };
    if constexpr (PARITY == 0) {
#pragma unroll
        for (int q = 0; q < 9; ++q) put(idx * B, (unsigned)opp(q) * plane * B, q, g[q], bv(q, 0, 0));
        if (OPEN && x == W - 2) { // reversed copy into the outlet ghost (lbm_core4.cu:114-116)
#pragma unroll
            for (int q = 0; q < 9; ++q) put((idx + 1) * B, (unsigned)opp(q) * plane * B, q, g[q], bv(q, 1, 0));
        }
    } else {
        // Bounce-back writes this cell's slot opp(q), otherwise the target's slot q (again a select).  The
        // populations leaving through the open inlet (c_qx < 0) are parked in this cell's slot opp(q): those
        // are the canonical inlet slots q = 1, 5, 8, which no other cell reads or writes in this step and which
        // the next step's Zou-He rebuilds without reading.
        const unsigned idxB = idx * B, pB = plane * B;
        const unsigned idx2 = idxB + 2u * pB;
        put(idxB, 0, 0, g[0], bv(0, 0, 0));
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const int o = opp(q);
            const bool back = solid[q] || (OPEN && inlet && cx(q) < 0);
            // q < o: slot q at the target = nbB[q] + q pB, slot o here = idx2 + q pB
            // q > o: slot q at the target = nb2[q] + o pB, slot o here = idxB + o pB
            const unsigned off = q < o ? (back ? idx2 : nbB[q]) : (back ? idxB : nb2[q]);
            // bounce-back: this row's slot that holds direction o next; else the target row's, direction q
            put(off, (unsigned)(q < o ? q : o) * pB, q, g[q], RB ? (back ? bv(o, 0, 0) : bv(q, cx(q), cy(q))) : 0.0f);
        }
        if (OPEN && x == W - 2) {
            // the ghost (W-1, y) is a copy of this cell: stream its q = 3, 6, 7 into column W-2
            put(idx * B, 3u * plane * B, 3, g[3], bv(3, 0, 0));
            if (!solid[2]) put(nb[2] * B, 6u * plane * B, 6, g[6], bv(6, 0, 1));
            if (!solid[4]) put(nb[4] * B, 7u * plane * B, 7, g[7], bv(7, 0, -1));
        }
    }
    if constexpr (MACROS) {
        orho[idx] = mac[0];
        oux[idx] = mac[1];
        ouy[idx] = mac[2];
    }
}

// One AA step of the thread's cell of the TILE_X x TILE_Y tile (bx, by), in the interior form when the whole
// tile satisfies aa_cell's INTERIOR condition (a workgroup-uniform branch).
template <int MODEL, int FMT, int PARITY, bool PERIODIC, bool MACROS>
__device__ __forceinline__ void aa_tile_cell(const KPar &P, const uint8_t *__restrict__ mask, void *__restrict__ f,
                                             const float *__restrict__ inlet_u, float *__restrict__ oux,
                                             float *__restrict__ ouy, float *__restrict__ orho, int bx, int by)
{
    const int x0 = bx * TILE_X, y0 = by * TILE_Y;
    const int x = x0 + (int)threadIdx.x, y = y0 + (int)threadIdx.y;
    if constexpr (row_based<FMT>()) {
        // tiles whose cells or neighbours meet the base window take the per-cell base form (general tile form)
        if (meets_window(P, x0 - 1, x0 + TILE_X, y0 - 1, y0 + TILE_Y)) {
            aa_cell<MODEL, FMT, PARITY, PERIODIC, MACROS, false, false, true>(P, mask, f, inlet_u, oux, ouy, orho, x, y);
            return;
        }
    }
    if (x0 >= 1 && x0 + TILE_X + 2 <= P.W && y0 >= 1 && y0 + TILE_Y + 1 <= P.H)
        aa_cell<MODEL, FMT, PARITY, PERIODIC, MACROS, false, true>(P, mask, f, inlet_u, oux, ouy, orho, x, y);
    else
        aa_cell<MODEL, FMT, PARITY, PERIODIC, MACROS>(P, mask, f, inlet_u, oux, ouy, orho, x, y);
}

} // namespace lbm
