// Temporally blocked k-step kernel (lbm_step_fused): k AA collide-stream steps per launch, bit-identical
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
// into registers.
#include "lbm_common.h"

namespace lbm {

namespace fused {

constexpr int MAXK = 8;
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

template <int MODEL, bool PERIODIC, int K, int P0, int LW>
__global__ __launch_bounds__((Sched<K, P0, LW>::NT)) void k_fused(KPar P, const uint8_t *__restrict__ mask,
                                                                  const float *__restrict__ fin,
                                                                  float *__restrict__ fout,
                                                                  const float *__restrict__ inlet_u, int CH, int macros,
                                                                  float *__restrict__ oux, float *__restrict__ ouy,
                                                                  float *__restrict__ orho)
{
    using S = Sched<K, P0, LW>;
    constexpr int R = S::R;
    constexpr int NT = S::NT;
    constexpr int CW = S::CW;
    constexpr int HALO = S::HALO;
    constexpr bool FINAL_REVERSED = ((P0 + K) & 1) != 0;
    __shared__ float sf[9][R][LW];
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

    // ---- level-0 rows travel HBM -> registers (PF rows ahead) -> LDS ring ------------------------
    float pre[PF][S::NLOAD];
    uint8_t pfl[PF];
    auto issue_load = [&](int j, float(&buf)[S::NLOAD], uint8_t &preflag) { // global row ybeg + j
        const int y = ybeg + j;
        bool rowin = PERIODIC ? true : (y >= 0 && y < H);
        const int gy = PERIODIC ? ((y % H) + H) % H : y;
#pragma unroll
        for (int n = 0; n < S::NLOAD; ++n) {
            const int e = t + n * NT;
            buf[n] = 0.0f;
            if (e < 9 * LW && rowin && j <= yend - ybeg) {
                const int q = e / LW, c = e - q * LW;
                int x = xs + c;
                bool colin = true;
                if (PERIODIC) {
                    x = ((x % W) + W) % W;
                } else {
                    colin = x >= 0 && x < W;
                }
                if (colin) buf[n] = fin[q * plane + gy * W + x];
            }
        }
        if (t < LW) {
            uint8_t fl = 0;
            int x = xs + t;
            if (PERIODIC) {
                x = ((x % W) + W) % W;
                if (mask[gy * W + x]) fl = F_SOLID;
                else fl = F_COMPUTE;
            } else if (rowin && x >= 0 && x < W) {
                const bool solid = mask[gy * W + x] != 0;
                if (solid) fl |= F_SOLID;
                if (gy >= 1 && gy <= H - 2 && x <= W - 2) {
                    if (x == 0) fl |= F_COMPUTE | F_INLET;
                    else if (!solid) fl |= F_COMPUTE;
                    if (x == W - 2) fl |= F_OUTADJ;
                }
            }
            preflag = fl;
        }
    };
    auto commit_load = [&](int j, const float(&buf)[S::NLOAD], uint8_t preflag) {
        const int slot = j % R;
#pragma unroll
        for (int n = 0; n < S::NLOAD; ++n) {
            const int e = t + n * NT;
            if (e < 9 * LW) {
                const int q = e / LW, c = e - q * LW;
                sf[q][slot][c] = buf[n];
            }
        }
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
        const int s = t / LW, c = t - (t / LW) * LW;
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
                        for (int q = 0; q < 9; ++q) g[q] = sf[q][sl][c];
                        collide<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac);
#pragma unroll
                        for (int q = 0; q < 9; ++q) sf[opp(q)][sl][c] = g[q];
                        if (!PERIODIC && (fl & F_OUTADJ) && c + 1 < LW) {
#pragma unroll
                            for (int q = 0; q < 9; ++q) sf[opp(q)][sl][c + 1] = g[q];
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
                        g[0] = sf[0][sl][c];
#pragma unroll
                        for (int q = 1; q < 9; ++q) {
                            const int o = opp(q);
                            if (!PERIODIC && inlet && cx(q) > 0)
                                g[q] = 0.0f;
                            else
                                g[q] = solid[o] ? sf[q][sl][c] : sf[o][rs[o]][cs[o]];
                        }
                        collide<MODEL>(P, g, inlet, inlet ? inlet_u[gy] : 0.0f, mac);
                        sf[0][sl][c] = g[0];
#pragma unroll
                        for (int q = 1; q < 9; ++q) {
                            if (!PERIODIC && inlet && cx(q) < 0) continue;
                            if (solid[q])
                                sf[opp(q)][sl][c] = g[q];
                            else
                                sf[q][rs[q]][cs[q]] = g[q];
                        }
                        if (!PERIODIC && (fl & F_OUTADJ)) {
                            sf[3][sl][c] = g[3];
                            if (!solid[2]) sf[6][su][c] = g[6];
                            if (!solid[4]) sf[7][sd][c] = g[7];
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
            for (int e = t; e < 9 * CW; e += NT) {
                const int q = e / CW, cc = e - q * CW;
                const int x = xs + HALO + cc;
                if (x >= W) continue;
                // The outlet ghost W-1 is part of cell W-2's state (its post-collision copy after an even
                // step), so the workgroup that owns W-2 writes it: the owner of W-1 may not have kept
                // W-2 valid through the last step.  A canonical ghost is never read and only written if
                // present.
                if (!PERIODIC && x == W - 1) continue;
                const float v = sf[q][sl][HALO + cc];
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

template <int MODEL, bool PERIODIC, int K, int P0, int LW>
static void launch_fused_t(const KPar &P, hipStream_t st, const uint8_t *mask, const float *fin, float *fout,
                           const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
    using S = Sched<K, P0, LW>;
    dim3 grid((P.W + S::CW - 1) / S::CW, (P.H + CH - 1) / CH);
    hipLaunchKernelGGL((k_fused<MODEL, PERIODIC, K, P0, LW>), grid, dim3(S::NT), 0, st, P, mask, fin, fout, inlet_u, CH,
                       rho != nullptr, ux, uy, rho);
}

template <int MODEL, bool PERIODIC, int K, int LW>
static void launch_fused_p(const KPar &P, int parity, hipStream_t st, const uint8_t *mask, const float *fin, float *fout,
                           const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
    if (parity)
        launch_fused_t<MODEL, PERIODIC, K, 1, LW>(P, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
    else
        launch_fused_t<MODEL, PERIODIC, K, 0, LW>(P, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
}

template <int MODEL, bool PERIODIC, int K>
static void launch_fused_w(const KPar &P, int lw, int parity, hipStream_t st, const uint8_t *mask, const float *fin,
                           float *fout, const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
    if (lw == 64)
        launch_fused_p<MODEL, PERIODIC, K, 64>(P, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
    else
        launch_fused_p<MODEL, PERIODIC, K, 128>(P, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho);
}

template <int MODEL, bool PERIODIC>
static int launch_fused_k(const KPar &P, int k, int lw, int parity, hipStream_t st, const uint8_t *mask,
                          const float *fin, float *fout, const float *inlet_u, int CH, float *ux, float *uy, float *rho)
{
#define LBM_FUSED_CASE(KK)                                                                                              \
    case KK: launch_fused_w<MODEL, PERIODIC, KK>(P, lw, parity, st, mask, fin, fout, inlet_u, CH, ux, uy, rho); break;
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

// Default launch configuration: 128-column strips, and the tallest chunk that still gives at least 512
// workgroups (two per CU on 256 CUs) so the k-row warm-up at each chunk end stays small.
static void default_cfg(int W, int H, int k, lbm_fused_cfg *c)
{
    c->strip = 128;
    const int CW = 128 - 2 * (k + 1);
    const int strips = (W + CW - 1) / CW;
    int ch = 256;
    while (ch > 16 && (long long)strips * ((H + ch - 1) / ch) < 512) ch /= 2;
    c->chunk = ch < H ? ch : H;
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_fused_max_k(void) { return fused::MAXK; }

int lbm_fused_default_cfg(const lbm_params *p, int k, lbm_fused_cfg *cfg)
{
    if (!p || !cfg) return fail(LBM_ERR_ARG, "lbm_fused_default_cfg: NULL pointer");
    default_cfg(p->W, p->H, k, cfg);
    return LBM_OK;
}

int lbm_step_fused_cfg(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out,
                       const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                       float *rho, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P);
    if (rc) return rc;
    if (!mask || !f_now || !f_out) return fail(LBM_ERR_ARG, "lbm_step_fused: mask, f_now, f_out must be non-NULL");
    {
        const char *a = reinterpret_cast<const char *>(f_now), *b = reinterpret_cast<const char *>(f_out);
        const size_t n = sizeof(float) * 9 * (size_t)P.plane;
        if (a < b + n && b < a + n) return fail(LBM_ERR_ARG, "lbm_step_fused: f_now and f_out overlap");
    }
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_step_fused: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_step_fused: parity must be 0 or 1");
    if ((ux == nullptr) != (rho == nullptr) || (uy == nullptr) != (rho == nullptr))
        return fail(LBM_ERR_ARG, "lbm_step_fused: ux, uy, rho must be all NULL or all non-NULL");
    if (k < 1 || k > fused::MAXK) return fail(LBM_ERR_ARG, "lbm_step_fused: k=%d outside [1, %d]", k, fused::MAXK);
    lbm_fused_cfg c;
    if (cfg)
        c = *cfg;
    else
        default_cfg(P.W, P.H, k, &c);
    if (c.strip != 64 && c.strip != 128) return fail(LBM_ERR_ARG, "lbm_step_fused: strip must be 64 or 128");
    if (c.chunk < 1) return fail(LBM_ERR_ARG, "lbm_step_fused: chunk must be >= 1");
    if (c.strip - 2 * (k + 1) < 16) return fail(LBM_ERR_ARG, "lbm_step_fused: strip %d too narrow for k=%d", c.strip, k);
    if (c.chunk > P.H) c.chunk = P.H;
    hipStream_t st = to_hip(stream);
    const bool per = p->bc == LBM_BC_PERIODIC;
    if (p->model == LBM_BGK)
        rc = per ? launch_fused_k<MODEL_BGK, true>(P, k, c.strip, parity, st, mask, f_now, f_out, inlet_u, c.chunk, ux,
                                                   uy, rho)
                 : launch_fused_k<MODEL_BGK, false>(P, k, c.strip, parity, st, mask, f_now, f_out, inlet_u, c.chunk, ux,
                                                    uy, rho);
    else
        rc = per ? launch_fused_k<MODEL_MRT, true>(P, k, c.strip, parity, st, mask, f_now, f_out, inlet_u, c.chunk, ux,
                                                   uy, rho)
                 : launch_fused_k<MODEL_MRT, false>(P, k, c.strip, parity, st, mask, f_now, f_out, inlet_u, c.chunk, ux,
                                                    uy, rho);
    if (rc) return rc;
    return check_launch("lbm_step_fused");
}

int lbm_step_fused(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out, const float *inlet_u,
                   int parity, int k, float *ux, float *uy, float *rho, lbm_stream_t stream)
{
    return lbm_step_fused_cfg(p, mask, f_now, f_out, inlet_u, parity, k, nullptr, ux, uy, rho, stream);
}

} // extern "C"
