// Single-step gfx950 kernels of liblbm and their C-ABI entry points:
//   lbm_step_aa    in-place AA step (the plain-method hot path)
//   lbm_step_pull  reference-compatible two-buffer pull step (lbm_core4.cu:7-212)
//   lbm_macro      rho/u extraction (lbm_core_smemaa_endfused.cu:377-422)
//   lbm_visualize  vorticity colour map (lbm_core4.cu:214-255)
//   lbm_diagnostics deterministic FP64 mass / momentum / momentum-exchange force
//
// Layout: SoA planes f[q*plane + y*W + x]; a 256-thread block covers 128 x 2 cells, so each wave64 reads
// and writes one 256-byte row segment per plane (coalesced); blocks are renumbered per XCD: bands of rows
// (xcd_tile) while the state fits the MALL, tile rows dealt round-robin (xcd_row_tile) beyond it.
// Odd (B) steps read and write the same 9 locations per cell (x - c_q / x + c_q pairs), which is what
// makes the single-buffer update race free.
#include "lbm_common.h"
#include "lbm_aa_cell.h"

#include <cstring>
#include <mutex>

namespace lbm {

static thread_local char g_err[512];

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail((int)e, "%s: launch failed: %s", what, hipGetErrorString(e));
    return LBM_OK;
}

int make_kpar(const lbm_params *p, KPar *k, unsigned pop_bytes)
{
    if (!p) return fail(LBM_ERR_ARG, "params is NULL");
    if (p->W < 3 || p->H < 3) return fail(LBM_ERR_SIZE, "lattice %dx%d too small (need W, H >= 3)", p->W, p->H);
    // element offsets q*plane + idx are kept in signed 32 bits
    if ((long long)p->W * p->H * 9 >= (1LL << 31))
        return fail(LBM_ERR_SIZE, "lattice %dx%d too large (9*W*H must be < 2^31)", p->W, p->H);
    // buffer resources: the record size 9*plane*bytes and every byte offset are unsigned 32-bit
    if ((long long)p->W * p->H * 9 * pop_bytes >= (1LL << 32))
        return fail(LBM_ERR_SIZE, "lattice %dx%d too large for %u-byte populations (9*W*H*%u must be < 2^32)", p->W,
                    p->H, pop_bytes, pop_bytes);
    if (p->model != LBM_MRT_LES && p->model != LBM_BGK) return fail(LBM_ERR_ARG, "unknown model %d", p->model);
    if (p->bc != LBM_BC_CHANNEL && p->bc != LBM_BC_PERIODIC) return fail(LBM_ERR_ARG, "unknown bc %d", p->bc);
    if (!(p->tau_inv > 0.0f && p->tau_inv < 2.0f)) return fail(LBM_ERR_ARG, "tau_inv %g outside (0,2)", p->tau_inv);
    if (!(p->cs >= 0.0f)) return fail(LBM_ERR_ARG, "cs %g must be >= 0", p->cs);
    k->W = p->W;
    k->H = p->H;
    k->plane = p->W * p->H;
    k->les = (p->model == LBM_MRT_LES && p->cs != 0.0f) ? 1 : 0;
    k->clamp = p->clamp ? 1 : 0;
    k->tau_inv = p->tau_inv;
    volatile float one = 1.0f; // keep the host division in FP32, as the device would do it
    k->tau0 = one / p->tau_inv;
    k->tau0sq = k->tau0 * k->tau0;
    k->c18cs2 = (18.0f * p->cs) * p->cs;
    k->s_e = p->s_e;
    k->s_eps = p->s_eps;
    k->s_q = p->s_q;
    if (p->sums != 0 && p->sums != 1) return fail(LBM_ERR_ARG, "sums %d must be 0 or 1", p->sums);
    k->sums = p->sums;
    if (p->fast != 0 && p->fast != 1) return fail(LBM_ERR_ARG, "fast %d must be 0 or 1", p->fast);
    if (p->fast && (p->sums || p->model != LBM_MRT_LES))
        return fail(LBM_ERR_UNSUPPORTED, "fast arithmetic is for the MRT model without compensated sums");
    k->fast = p->fast;
    if (p->row_base && (reinterpret_cast<uintptr_t>(p->row_base) & 15u))
        return fail(LBM_ERR_ARG, "row_base must be 16-byte aligned");
    k->base = p->row_base;
    k->wx0 = k->wy0 = k->ww = k->wh = 0;
    k->wvel = nullptr;
    k->wtx0 = k->wty0 = k->wtw = k->wth = 0;
    if (p->window) {
        const lbm_base_window &w = *p->window;
        if (!p->row_base) return fail(LBM_ERR_ARG, "a base window needs row_base (the base outside the window)");
        if (!w.vel || (reinterpret_cast<uintptr_t>(w.vel) & 3u))
            return fail(LBM_ERR_ARG, "the window's velocity table must be non-NULL and 4-byte aligned");
        // not x = W - 2: the two-step kernel's outlet ghost is a copy of that column's words, valid only when
        // both columns share their rows' base
        if (w.w < 1 || w.h < 1 || w.x0 < 1 || w.y0 < 1 || (long long)w.x0 + w.w > p->W - 2 ||
            (long long)w.y0 + w.h > p->H - 1)
            return fail(LBM_ERR_ARG, "base window [%d, %d) x [%d, %d) must lie in 1 <= x <= W - 3, 1 <= y <= H - 2",
                        w.x0, w.x0 + w.w, w.y0, w.y0 + w.h);
        k->wx0 = w.x0;
        k->wy0 = w.y0;
        k->ww = w.w;
        k->wh = w.h;
        k->wvel = reinterpret_cast<const uint32_t *>(w.vel);
    }
    // measured on the cylinder case: the banded order is 1-2 % faster while the state fits the MALL (226 MB),
    // the row order 5-10 % faster beyond it (302 MB and up)
    k->rowmap = (long long)p->W * p->H * 9 * pop_bytes > (256LL << 20) ? 1 : 0;
    return LBM_OK;
}

// ------------------------------------------------------------------------------------------------------
// in-place AA step
// ------------------------------------------------------------------------------------------------------

// Index of the neighbour x + c_q (wrapped when periodic; only called for in-domain neighbours otherwise).
template <bool PERIODIC>
__device__ __forceinline__ int nbr_index(const KPar &P, int x, int y, int q)
{
    int nx = x + cx(q), ny = y + cy(q);
    if constexpr (PERIODIC) {
        nx = nx < 0 ? nx + P.W : (nx >= P.W ? nx - P.W : nx);
        ny = ny < 0 ? ny + P.H : (ny >= P.H ? ny - P.H : ny);
    }
    return ny * P.W + nx;
}

template <int MODEL, int FMT, int PARITY, bool PERIODIC, bool MACROS>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_aa(KPar P, const uint8_t *__restrict__ mask, void *__restrict__ f,
                                                        const float *__restrict__ inlet_u, float *__restrict__ oux,
                                                        float *__restrict__ ouy, float *__restrict__ orho)
{
    int bx, by;
    if (P.rowmap)
        xcd_row_tile(bx, by);
    else
        xcd_tile(bx, by);
    aa_tile_cell<MODEL, FMT, PARITY, PERIODIC, MACROS>(P, mask, f, inlet_u, oux, ouy, orho, bx, by);
}

template <int MODEL, int FMT, int PARITY, bool PERIODIC, bool MACROS>
static void launch_aa(const KPar &P, dim3 grid, hipStream_t s, const uint8_t *mask, void *f, const float *inlet_u,
                      float *ux, float *uy, float *rho)
{
    hipLaunchKernelGGL((k_aa<MODEL, FMT, PARITY, PERIODIC, MACROS>), grid, dim3(TILE_X, TILE_Y), 0, s, P, mask, f,
                       inlet_u, ux, uy, rho);
}

template <int MODEL, int FMT, int PARITY, bool PERIODIC>
static void dispatch_aa_m(const KPar &P, dim3 grid, hipStream_t s, const uint8_t *mask, void *f, const float *inlet_u,
                          float *ux, float *uy, float *rho)
{
    if (rho)
        launch_aa<MODEL, FMT, PARITY, PERIODIC, true>(P, grid, s, mask, f, inlet_u, ux, uy, rho);
    else
        launch_aa<MODEL, FMT, PARITY, PERIODIC, false>(P, grid, s, mask, f, inlet_u, ux, uy, rho);
}

template <int MODEL, int FMT>
static void dispatch_aa(const KPar &P, int bc, int parity, dim3 grid, hipStream_t s, const uint8_t *mask, void *f,
                        const float *inlet_u, float *ux, float *uy, float *rho)
{
    if (bc == LBM_BC_PERIODIC) {
        if (parity)
            dispatch_aa_m<MODEL, FMT, 1, true>(P, grid, s, mask, f, inlet_u, ux, uy, rho);
        else
            dispatch_aa_m<MODEL, FMT, 0, true>(P, grid, s, mask, f, inlet_u, ux, uy, rho);
    } else {
        if (parity)
            dispatch_aa_m<MODEL, FMT, 1, false>(P, grid, s, mask, f, inlet_u, ux, uy, rho);
        else
            dispatch_aa_m<MODEL, FMT, 0, false>(P, grid, s, mask, f, inlet_u, ux, uy, rho);
    }
}

template <int FMT>
static void dispatch_aa_fmt(const KPar &P, int model, int bc, int parity, dim3 grid, hipStream_t s,
                            const uint8_t *mask, void *f, const float *inlet_u, float *ux, float *uy, float *rho)
{
    if (model == LBM_BGK)
        dispatch_aa<MODEL_BGK, FMT>(P, bc, parity, grid, s, mask, f, inlet_u, ux, uy, rho);
    else if (P.fast)
        dispatch_aa<MODEL_MRT | MODEL_FAST, FMT>(P, bc, parity, grid, s, mask, f, inlet_u, ux, uy, rho);
    else if (P.sums)
        dispatch_aa<MODEL_MRT_SUMS, FMT>(P, bc, parity, grid, s, mask, f, inlet_u, ux, uy, rho);
    else
        dispatch_aa<MODEL_MRT, FMT>(P, bc, parity, grid, s, mask, f, inlet_u, ux, uy, rho);
}

// ------------------------------------------------------------------------------------------------------
// reference-compatible pull step (lbm_core4.cu:120-211)
// ------------------------------------------------------------------------------------------------------

template <int MODEL, bool PERIODIC>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_pull(KPar P, const uint8_t *__restrict__ mask,
                                                          const float *__restrict__ f_now, float *__restrict__ f_out,
                                                          const float *__restrict__ inlet_u, float *__restrict__ oux,
                                                          float *__restrict__ ouy, float *__restrict__ orho)
{
    int bx, by;
    xcd_tile(bx, by);
    const int x = bx * TILE_X + threadIdx.x;
    const int y = by * TILE_Y + threadIdx.y;
    if (x >= P.W || y >= P.H) return;
    const int W = P.W, plane = P.plane;
    const int idx = y * W + x;
    bool inlet = false;
    if constexpr (!PERIODIC) {
        if (y == 0 || y == P.H - 1 || x == W - 1) return;
        inlet = x == 0;
        if (!inlet && mask[idx]) return;
    } else {
        if (mask[idx]) return;
    }
    float g[9], mac[3];
    g[0] = f_now[idx];
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        if (!PERIODIC && inlet && cx(q) > 0) {
            g[q] = 0.0f; // lbm_core4.cu:36-37
        } else {
            const int s = nbr_index<PERIODIC>(P, x, y, opp(q)); // x - c_q
            g[q] = mask[s] ? f_now[opp(q) * plane + idx] : f_now[q * plane + s];
        }
    }
    collide<MODEL>(P, g, inlet, inlet ? inlet_u[y] : 0.0f, mac);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        f_out[q * plane + idx] = g[q];
        if (!PERIODIC && x == W - 2) f_out[q * plane + idx + 1] = g[q];
    }
    orho[idx] = mac[0];
    oux[idx] = mac[1];
    ouy[idx] = mac[2];
}

// ------------------------------------------------------------------------------------------------------
// macro extraction (lbm_core_smemaa_endfused.cu:377-422), both parities
// ------------------------------------------------------------------------------------------------------

template <bool PERIODIC, int FMT>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_macro(KPar P, const uint8_t *__restrict__ mask,
                                                           const typename Conv<FMT>::T *__restrict__ f,
                                                           const float *__restrict__ inlet_u, int parity,
                                                           float *__restrict__ oux, float *__restrict__ ouy,
                                                           float *__restrict__ orho)
{
    const int x = blockIdx.x * TILE_X + threadIdx.x;
    const int y = blockIdx.y * TILE_Y + threadIdx.y;
    if (x >= P.W || y >= P.H) return;
    const int idx = y * P.W + x;
    bool inlet = false;
    if constexpr (!PERIODIC) {
        if (y == 0 || y == P.H - 1 || x == P.W - 1) return;
        inlet = x == 0;
        if (!inlet && mask[idx]) return;
    } else {
        if (mask[idx]) return;
    }
    float g[9];
    if constexpr (row_based<FMT>()) {
        // the slot holding direction q lies on this row: base value d[q] (a wave covers part of one row), or the
        // cell's own in the base window
        float d[9];
        base_row_s(P.base, y, d);
        if (in_window(P, x, y)) {
            float wx, wy;
            unpack_vel(window_vel(P, x, y), wx, wy);
            const float us = vel_usq15(wx, wy);
#pragma unroll
            for (int q = 0; q < 9; ++q) d[q] = vel_base(q, wx, wy, us);
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = Conv<FMT>::popb(f[(parity ? opp(q) : q) * P.plane + idx], q, d[q]);
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = Conv<FMT>::pop(f[(parity ? opp(q) : q) * P.plane + idx], q);
    }
    float rho, ux, uy;
    if (inlet && parity == 0) {
        const float u = inlet_u[y];
        rho = __fdiv_rn(fma_(2.0f, (g[3] + g[6]) + g[7], (g[0] + g[2]) + g[4]), 1.0f - u);
        ux = u;
        uy = 0.0f;
    } else {
        const float s13 = g[1] + g[3], s24 = g[2] + g[4], s57 = g[5] + g[7], s68 = g[6] + g[8];
        const float a13 = g[1] - g[3], a24 = g[2] - g[4], a57 = g[5] - g[7], a86 = g[8] - g[6];
        rho = (g[0] + (s13 + s24)) + (s57 + s68);
        ux = __fdiv_rn(a13 + (a57 + a86), rho);
        uy = __fdiv_rn(a24 + (a57 - a86), rho);
    }
    orho[idx] = rho;
    oux[idx] = ux;
    ouy[idx] = uy;
}

// ------------------------------------------------------------------------------------------------------
// vorticity colour map (lbm_core4.cu:214-255)
// ------------------------------------------------------------------------------------------------------

__global__ __launch_bounds__(TILE_X *TILE_Y) void k_visualize(int W, int H, const float *__restrict__ ux,
                                                               const float *__restrict__ uy,
                                                               const uint8_t *__restrict__ mask,
                                                               uint8_t *__restrict__ image, float vort_scale)
{
    const int x = blockIdx.x * TILE_X + threadIdx.x;
    const int y = blockIdx.y * TILE_Y + threadIdx.y;
    if (x <= 0 || x >= W - 1 || y <= 0 || y >= H - 1) return;
    const int pid = y * W + x;
    image[pid * 4 + 3] = 255;
    if (mask[pid]) {
        image[pid * 4 + 0] = 0;
        image[pid * 4 + 1] = 0;
        image[pid * 4 + 2] = 0;
        return;
    }
    const float vort = ((uy[pid + 1] - uy[pid - 1]) - (ux[pid + W] - ux[pid - W])) * vort_scale;
    const float r = fminf(fmaxf(1.0f + vort, 0.0f), 1.0f);
    const float g = fminf(fmaxf(1.0f - fabsf(vort), 0.0f), 1.0f);
    const float b = fminf(fmaxf(1.0f - vort, 0.0f), 1.0f);
    image[pid * 4 + 0] = (uint8_t)(r * 255.0f);
    image[pid * 4 + 1] = (uint8_t)(g * 255.0f);
    image[pid * 4 + 2] = (uint8_t)(b * 255.0f);
}

// ------------------------------------------------------------------------------------------------------
// format conversion of a whole state (populations f <-> deviations h = f - w)
// ------------------------------------------------------------------------------------------------------

template <int SRC, int DST>
__global__ __launch_bounds__(256) void k_convert(unsigned plane, const typename Conv<SRC>::T *__restrict__ src,
                                                  typename Conv<DST>::T *__restrict__ dst)
{
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= plane) return;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        dst[(unsigned)q * plane + i] = Conv<DST>::from_pop(Conv<SRC>::pop(src[(unsigned)q * plane + i], q), q);
}

// Conversion with a per-row base on either side (row-based 16-bit formats): slot p of a cell on row y holds direction
// p (parity 0) or opp(p) (parity 1), whose base value on that row is applied.
template <int SRC, int DST>
__global__ __launch_bounds__(256) void k_convert_rb(KPar P, int parity, const typename Conv<SRC>::T *__restrict__ src,
                                                     typename Conv<DST>::T *__restrict__ dst)
{
    const unsigned plane = (unsigned)P.plane;
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= plane) return;
    const int y = (int)(i / (unsigned)P.W), x = (int)(i - (unsigned)y * (unsigned)P.W);
    float bt[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) bt[q] = P.base[ROWB * y + q];
    if (in_window(P, x, y)) {
        float wx, wy;
        unpack_vel(window_vel(P, x, y), wx, wy);
        const float us = vel_usq15(wx, wy);
#pragma unroll
        for (int q = 0; q < 9; ++q) bt[q] = vel_base(q, wx, wy, us);
    }
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int q = parity ? opp(p) : p;
        const auto v = src[(unsigned)p * plane + i];
        float f;
        if constexpr (row_based<SRC>())
            f = Conv<SRC>::popb(v, q, bt[q]);
        else
            f = Conv<SRC>::pop(v, q);
        if constexpr (row_based<DST>())
            dst[(unsigned)p * plane + i] = Conv<DST>::from_popb(f, q, bt[q]);
        else
            dst[(unsigned)p * plane + i] = Conv<DST>::from_pop(f, q);
    }
}

// Base window re-derivation: per window cell, decode its nine words with the current base, take the cell's velocity
// (0 where rho is outside (0.5, 2) or a component is not below 1 in magnitude), rounded to IEEE half, as its new
// base velocity, store it to nv and re-encode the words against the base values it gives (vel_base).  Each cell
// touches only its own slots (in either parity slot q of a cell holds one of that cell's populations), so cells are
// independent.
template <int FMT>
__global__ __launch_bounds__(256) void k_rebase(KPar P, int parity, uint16_t *__restrict__ h, uint32_t *__restrict__ nv)
{
    using C = Conv<FMT>;
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= (unsigned)(P.ww * P.wh)) return;
    const int x = P.wx0 + (int)(i % (unsigned)P.ww), y = P.wy0 + (int)(i / (unsigned)P.ww);
    const unsigned idx = (unsigned)(y * P.W + x), plane = (unsigned)P.plane;
    float ox, oy;
    unpack_vel(P.wvel[i], ox, oy);
    const float ous = vel_usq15(ox, oy);
    float f[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
        f[q] = C::popb(h[(unsigned)(parity ? opp(q) : q) * plane + idx], q, vel_base(q, ox, oy, ous));
    const float rho = (((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7]))) + f[8];
    const float jx = ((f[1] - f[3]) + (f[5] - f[6])) + (f[8] - f[7]);
    const float jy = ((f[2] - f[4]) + (f[5] + f[6])) - (f[7] + f[8]);
    const float ux = __fdiv_rn(jx, rho), uy = __fdiv_rn(jy, rho);
    const bool ok = rho > 0.5f && rho < 2.0f && __builtin_fabsf(ux) < 1.0f && __builtin_fabsf(uy) < 1.0f;
    const _Float16 hx = ok ? (_Float16)ux : (_Float16)0.0f, hy = ok ? (_Float16)uy : (_Float16)0.0f;
    const uint32_t packed = (uint32_t)__builtin_bit_cast(uint16_t, hx) | ((uint32_t)__builtin_bit_cast(uint16_t, hy) << 16);
    nv[i] = packed;
    float nx, ny;
    unpack_vel(packed, nx, ny);
    const float nus = vel_usq15(nx, ny);
#pragma unroll
    for (int q = 0; q < 9; ++q) h[(unsigned)(parity ? opp(q) : q) * plane + idx] = C::from_popb(f[q], q, vel_base(q, nx, ny, nus));
}

// ------------------------------------------------------------------------------------------------------
// FP64 diagnostics: fixed grid, fixed-order tree reductions -> bitwise reproducible
// ------------------------------------------------------------------------------------------------------

constexpr int DIAG_BLOCKS = 1024;
constexpr int DIAG_THREADS = 256;
constexpr int DIAG_N = 5;

template <bool PERIODIC>
__global__ __launch_bounds__(DIAG_THREADS) void k_diag_partial(KPar P, const uint8_t *__restrict__ mask,
                                                                const float *__restrict__ f,
                                                                const float *__restrict__ inlet_u, int parity,
                                                                double *__restrict__ work)
{
    __shared__ double red[DIAG_N][DIAG_THREADS];
    double acc[DIAG_N] = {0, 0, 0, 0, 0};
    const int ncell = P.plane;
    for (int i = blockIdx.x * DIAG_THREADS + threadIdx.x; i < ncell; i += DIAG_BLOCKS * DIAG_THREADS) {
        const int y = i / P.W, x = i - y * P.W;
        bool inlet = false;
        if constexpr (!PERIODIC) {
            if (y == 0 || y == P.H - 1 || x == P.W - 1) continue;
            inlet = x == 0;
            if (!inlet && mask[i]) continue;
        } else {
            if (mask[i]) continue;
        }
        double g[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = (double)f[(parity ? opp(q) : q) * P.plane + i];
        if (inlet && parity == 0) {
            const double u = (double)inlet_u[y];
            const double rho = (g[0] + g[2] + g[4] + 2.0 * (g[3] + g[6] + g[7])) / (1.0 - u);
            acc[0] += rho;
            acc[1] += rho * u;
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                acc[0] += g[q];
                acc[1] += g[q] * cx(q);
                acc[2] += g[q] * cy(q);
            }
        }
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            int nx = x + cx(q), ny = y + cy(q);
            if constexpr (PERIODIC) {
                nx = nx < 0 ? nx + P.W : (nx >= P.W ? nx - P.W : nx);
                ny = ny < 0 ? ny + P.H : (ny >= P.H ? ny - P.H : ny);
            } else {
                if (nx < 0 || ny <= 0 || ny >= P.H - 1) continue; // outside, or a wall row
            }
            if (!mask[ny * P.W + nx]) continue;
            const double fs = (double)f[opp(q) * P.plane + i];
            acc[3] += 2.0 * cx(q) * fs;
            acc[4] += 2.0 * cy(q) * fs;
        }
    }
#pragma unroll
    for (int c = 0; c < DIAG_N; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = DIAG_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int c = 0; c < DIAG_N; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_N) work[blockIdx.x * DIAG_N + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(DIAG_THREADS) void k_diag_final(const double *__restrict__ work, double *__restrict__ out)
{
    __shared__ double red[DIAG_N][DIAG_THREADS];
    double acc[DIAG_N] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < DIAG_BLOCKS; b += DIAG_THREADS) {
#pragma unroll
        for (int c = 0; c < DIAG_N; ++c) acc[c] += work[b * DIAG_N + c];
    }
#pragma unroll
    for (int c = 0; c < DIAG_N; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = DIAG_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int c = 0; c < DIAG_N; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_N) out[threadIdx.x] = red[threadIdx.x][0];
}

static dim3 step_grid(const KPar &P) { return dim3((P.W + TILE_X - 1) / TILE_X, (P.H + TILE_Y - 1) / TILE_Y); }

// Scratch of lbm_diagnostics_noalloc: the per-workgroup FP64 partial sums of k_diag_partial's fixed grid
// (DIAG_BLOCKS x DIAG_N doubles, 40 KiB), held by every GPU the library runs on until the process ends.  Host
// code must not pass `scratch` itself to a kernel (the name is a host-side stand-in, not a device address):
// scratch_ptr() looks up each device's own copy once, with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[DIAG_BLOCKS * DIAG_N];
constexpr size_t SCRATCH_BYTES = sizeof(double) * DIAG_BLOCKS * DIAG_N;
constexpr int SCRATCH_MAX_DEVICES = 64;
static double *g_scratch[SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The current device's address of `scratch` (looked up on the device's first call), or NULL with the error set.
static double *scratch_ptr(int &rc)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        rc = fail((int)e, "lbm_diagnostics_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return nullptr;
    }
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        rc = fail(LBM_ERR_UNSUPPORTED, "lbm_diagnostics_noalloc: device %d beyond the %d supported", dev,
                  SCRATCH_MAX_DEVICES);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            rc = fail(e != hipSuccess ? (int)e : LBM_ERR_ARG, "lbm_diagnostics_noalloc: hipGetSymbolAddress: %s",
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    rc = LBM_OK;
    return g_scratch[dev];
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_abi_version(void) { return 5; }

const char *lbm_last_error(void) { return g_err; }

int lbm_sync(lbm_stream_t stream)
{
    hipError_t e = hipStreamSynchronize(to_hip(stream));
    if (e != hipSuccess) return fail((int)e, "stream synchronize: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_step_aa(const lbm_params *p, const uint8_t *mask, float *f, const float *inlet_u, int parity, float *ux,
                float *uy, float *rho, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 4);
    if (rc) return rc;
    if (!mask || !f) return fail(LBM_ERR_ARG, "lbm_step_aa: mask and f must be non-NULL");
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_step_aa: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_step_aa: parity must be 0 or 1");
    if ((ux == nullptr) != (rho == nullptr) || (uy == nullptr) != (rho == nullptr))
        return fail(LBM_ERR_ARG, "lbm_step_aa: ux, uy, rho must be all NULL or all non-NULL");
    dispatch_aa_fmt<FMT_F32>(P, p->model, p->bc, parity, step_grid(P), to_hip(stream), mask, f, inlet_u, ux, uy, rho);
    return check_launch("lbm_step_aa");
}

int lbm_step_aa_n(const lbm_params *p, const uint8_t *mask, float *f, const float *inlet_u, int parity, int nsteps,
                  lbm_stream_t stream)
{
    if (nsteps < 0) return fail(LBM_ERR_ARG, "lbm_step_aa_n: nsteps must be >= 0");
    for (int i = 0; i < nsteps; ++i) {
        int rc = lbm_step_aa(p, mask, f, inlet_u, (parity + i) & 1, nullptr, nullptr, nullptr, stream);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm_step_pull(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out, float *ux, float *uy,
                  float *rho, const float *inlet_u, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!mask || !f_now || !f_out || !ux || !uy || !rho)
        return fail(LBM_ERR_ARG, "lbm_step_pull: mask, f_now, f_out, ux, uy, rho must be non-NULL");
    if (P.fast) return fail(LBM_ERR_UNSUPPORTED, "lbm_step_pull: fast arithmetic is not built for the pull step");
    if (f_now == f_out) return fail(LBM_ERR_ARG, "lbm_step_pull: f_now and f_out must differ");
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_step_pull: channel mode needs inlet_u");
    const dim3 grid = step_grid(P), block(TILE_X, TILE_Y);
    hipStream_t s = to_hip(stream);
    const bool per = p->bc == LBM_BC_PERIODIC;
    if (p->model == LBM_BGK) {
        if (per)
            hipLaunchKernelGGL((k_pull<MODEL_BGK, true>), grid, block, 0, s, P, mask, f_now, f_out, inlet_u, ux, uy, rho);
        else
            hipLaunchKernelGGL((k_pull<MODEL_BGK, false>), grid, block, 0, s, P, mask, f_now, f_out, inlet_u, ux, uy, rho);
    } else if (P.sums) {
        if (per)
            hipLaunchKernelGGL((k_pull<MODEL_MRT_SUMS, true>), grid, block, 0, s, P, mask, f_now, f_out, inlet_u, ux, uy,
                               rho);
        else
            hipLaunchKernelGGL((k_pull<MODEL_MRT_SUMS, false>), grid, block, 0, s, P, mask, f_now, f_out, inlet_u, ux, uy,
                               rho);
    } else {
        if (per)
            hipLaunchKernelGGL((k_pull<MODEL_MRT, true>), grid, block, 0, s, P, mask, f_now, f_out, inlet_u, ux, uy, rho);
        else
            hipLaunchKernelGGL((k_pull<MODEL_MRT, false>), grid, block, 0, s, P, mask, f_now, f_out, inlet_u, ux, uy, rho);
    }
    return check_launch("lbm_step_pull");
}

int lbm_macro(const lbm_params *p, const uint8_t *mask, const float *f, const float *inlet_u, int parity, float *ux,
              float *uy, float *rho, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!mask || !f || !ux || !uy || !rho) return fail(LBM_ERR_ARG, "lbm_macro: NULL pointer");
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_macro: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_macro: parity must be 0 or 1");
    const dim3 grid = step_grid(P), block(TILE_X, TILE_Y);
    hipStream_t s = to_hip(stream);
    if (p->bc == LBM_BC_PERIODIC)
        hipLaunchKernelGGL((k_macro<true, FMT_F32>), grid, block, 0, s, P, mask, f, inlet_u, parity, ux, uy, rho);
    else
        hipLaunchKernelGGL((k_macro<false, FMT_F32>), grid, block, 0, s, P, mask, f, inlet_u, parity, ux, uy, rho);
    return check_launch("lbm_macro");
}

int lbm_visualize(const lbm_params *p, const float *ux, const float *uy, const uint8_t *mask, uint8_t *rgba,
                  float vort_scale, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!ux || !uy || !mask || !rgba) return fail(LBM_ERR_ARG, "lbm_visualize: NULL pointer");
    const dim3 grid = step_grid(P), block(TILE_X, TILE_Y);
    hipLaunchKernelGGL(k_visualize, grid, block, 0, to_hip(stream), P.W, P.H, ux, uy, mask, rgba, vort_scale);
    return check_launch("lbm_visualize");
}

int lbm_step_aa_h16(const lbm_params *p, const uint8_t *mask, uint16_t *h, int fmt, const float *inlet_u, int parity,
                    float *ux, float *uy, float *rho, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 2);
    if (rc) return rc;
    if (!mask || !h) return fail(LBM_ERR_ARG, "lbm_step_aa_h16: mask and h must be non-NULL");
    if (fmt != LBM_FMT_F16 && fmt != LBM_FMT_BF16) return fail(LBM_ERR_ARG, "lbm_step_aa_h16: bad format %d", fmt);
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_step_aa_h16: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_step_aa_h16: parity must be 0 or 1");
    if ((ux == nullptr) != (rho == nullptr) || (uy == nullptr) != (rho == nullptr))
        return fail(LBM_ERR_ARG, "lbm_step_aa_h16: ux, uy, rho must be all NULL or all non-NULL");
    const dim3 grid = step_grid(P);
    hipStream_t s = to_hip(stream);
    if (P.base) { // deviations from the per-row base
        if (fmt == LBM_FMT_F16)
            dispatch_aa_fmt<FMT_F16R>(P, p->model, p->bc, parity, grid, s, mask, h, inlet_u, ux, uy, rho);
        else
            dispatch_aa_fmt<FMT_BF16R>(P, p->model, p->bc, parity, grid, s, mask, h, inlet_u, ux, uy, rho);
    } else if (fmt == LBM_FMT_F16) {
        dispatch_aa_fmt<FMT_F16>(P, p->model, p->bc, parity, grid, s, mask, h, inlet_u, ux, uy, rho);
    } else {
        dispatch_aa_fmt<FMT_BF16>(P, p->model, p->bc, parity, grid, s, mask, h, inlet_u, ux, uy, rho);
    }
    return check_launch("lbm_step_aa_h16");
}

int lbm_step_aa_h16_n(const lbm_params *p, const uint8_t *mask, uint16_t *h, int fmt, const float *inlet_u, int parity,
                      int nsteps, lbm_stream_t stream)
{
    if (nsteps < 0) return fail(LBM_ERR_ARG, "lbm_step_aa_h16_n: nsteps must be >= 0");
    for (int i = 0; i < nsteps; ++i) {
        int rc = lbm_step_aa_h16(p, mask, h, fmt, inlet_u, (parity + i) & 1, nullptr, nullptr, nullptr, stream);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm_macro_h16(const lbm_params *p, const uint8_t *mask, const uint16_t *h, int fmt, const float *inlet_u,
                  int parity, float *ux, float *uy, float *rho, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!mask || !h || !ux || !uy || !rho) return fail(LBM_ERR_ARG, "lbm_macro_h16: NULL pointer");
    if (fmt != LBM_FMT_F16 && fmt != LBM_FMT_BF16) return fail(LBM_ERR_ARG, "lbm_macro_h16: bad format %d", fmt);
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_macro_h16: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_macro_h16: parity must be 0 or 1");
    const dim3 grid = step_grid(P), block(TILE_X, TILE_Y);
    hipStream_t s = to_hip(stream);
    const bool per = p->bc == LBM_BC_PERIODIC;
    if (P.base) {
        if (fmt == LBM_FMT_F16) {
            if (per)
                hipLaunchKernelGGL((k_macro<true, FMT_F16R>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy, rho);
            else
                hipLaunchKernelGGL((k_macro<false, FMT_F16R>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy,
                                   rho);
        } else {
            if (per)
                hipLaunchKernelGGL((k_macro<true, FMT_BF16R>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy,
                                   rho);
            else
                hipLaunchKernelGGL((k_macro<false, FMT_BF16R>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy,
                                   rho);
        }
    } else if (fmt == LBM_FMT_F16) {
        if (per)
            hipLaunchKernelGGL((k_macro<true, FMT_F16>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy, rho);
        else
            hipLaunchKernelGGL((k_macro<false, FMT_F16>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy, rho);
    } else {
        if (per)
            hipLaunchKernelGGL((k_macro<true, FMT_BF16>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy, rho);
        else
            hipLaunchKernelGGL((k_macro<false, FMT_BF16>), grid, block, 0, s, P, mask, h, inlet_u, parity, ux, uy, rho);
    }
    return check_launch("lbm_macro_h16");
}

int lbm_convert(const lbm_params *p, const void *src, int src_fmt, void *dst, int dst_fmt, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!src || !dst) return fail(LBM_ERR_ARG, "lbm_convert: NULL pointer");
    if (P.base)
        return fail(LBM_ERR_ARG, "lbm_convert: p->row_base is set; lbm_convert_state applies it (it needs the parity)");
    auto bytes = [](int f) { return f == LBM_FMT_F32 ? 4u : 2u; };
    if (src_fmt < 0 || src_fmt > 2 || dst_fmt < 0 || dst_fmt > 2) return fail(LBM_ERR_ARG, "lbm_convert: bad format");
    {
        const char *a = static_cast<const char *>(src), *b = static_cast<const char *>(dst);
        if (a < b + 9ull * P.plane * bytes(dst_fmt) && b < a + 9ull * P.plane * bytes(src_fmt))
            return fail(LBM_ERR_ARG, "lbm_convert: src and dst overlap");
    }
    const dim3 grid((P.plane + 255) / 256), block(256);
    hipStream_t s = to_hip(stream);
    const unsigned plane = (unsigned)P.plane;
#define LBM_CONV(S, D)                                                                                                  \
    if (src_fmt == S && dst_fmt == D) {                                                                                 \
        hipLaunchKernelGGL((k_convert<S, D>), grid, block, 0, s, plane,                                                 \
                           static_cast<const typename Conv<S>::T *>(src), static_cast<typename Conv<D>::T *>(dst));     \
        return check_launch("lbm_convert");                                                                             \
    }
    LBM_CONV(FMT_F32, FMT_F16)
    LBM_CONV(FMT_F32, FMT_BF16)
    LBM_CONV(FMT_F16, FMT_F32)
    LBM_CONV(FMT_BF16, FMT_F32)
    LBM_CONV(FMT_F16, FMT_BF16)
    LBM_CONV(FMT_BF16, FMT_F16)
    LBM_CONV(FMT_F32, FMT_F32)
#undef LBM_CONV
    return fail(LBM_ERR_UNSUPPORTED, "lbm_convert: %d -> %d not supported", src_fmt, dst_fmt);
}

int lbm_convert_state(const lbm_params *p, const void *src, int src_fmt, void *dst, int dst_fmt, int parity,
                      lbm_stream_t stream)
{
    if (!p) return fail(LBM_ERR_ARG, "params is NULL");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_convert_state: parity must be 0 or 1");
    if (!p->row_base) return lbm_convert(p, src, src_fmt, dst, dst_fmt, stream);
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!src || !dst) return fail(LBM_ERR_ARG, "lbm_convert_state: NULL pointer");
    auto bytes = [](int f) { return f == LBM_FMT_F32 ? 4u : 2u; };
    if (src_fmt < 0 || src_fmt > 2 || dst_fmt < 0 || dst_fmt > 2)
        return fail(LBM_ERR_ARG, "lbm_convert_state: bad format");
    {
        const char *a = static_cast<const char *>(src), *b = static_cast<const char *>(dst);
        if (a < b + 9ull * P.plane * bytes(dst_fmt) && b < a + 9ull * P.plane * bytes(src_fmt))
            return fail(LBM_ERR_ARG, "lbm_convert_state: src and dst overlap");
    }
    const dim3 grid((P.plane + 255) / 256), block(256);
    hipStream_t s = to_hip(stream);
    // the 16-bit sides are row-based
#define LBM_CONV(S, D, KS, KD)                                                                                          \
    if (src_fmt == S && dst_fmt == D) {                                                                                 \
        hipLaunchKernelGGL((k_convert_rb<KS, KD>), grid, block, 0, s, P, parity,                                       \
                           static_cast<const typename Conv<KS>::T *>(src), static_cast<typename Conv<KD>::T *>(dst));   \
        return check_launch("lbm_convert_state");                                                                       \
    }
    LBM_CONV(FMT_F32, FMT_F16, FMT_F32, FMT_F16R)
    LBM_CONV(FMT_F32, FMT_BF16, FMT_F32, FMT_BF16R)
    LBM_CONV(FMT_F16, FMT_F32, FMT_F16R, FMT_F32)
    LBM_CONV(FMT_BF16, FMT_F32, FMT_BF16R, FMT_F32)
    LBM_CONV(FMT_F16, FMT_BF16, FMT_F16R, FMT_BF16R)
    LBM_CONV(FMT_BF16, FMT_F16, FMT_BF16R, FMT_F16R)
#undef LBM_CONV
    // src_fmt == dst_fmt: the stored words themselves
    const hipError_t e = hipMemcpyAsync(dst, src, 9ull * P.plane * bytes(src_fmt), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return fail((int)e, "lbm_convert_state: copy failed: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_rebase_window(const lbm_params *p, uint16_t *h, int fmt, int parity, uint16_t *new_vel, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 2);
    if (rc) return rc;
    if (!p->window) return fail(LBM_ERR_ARG, "lbm_rebase_window: p->window is NULL");
    if (!h || !new_vel || (reinterpret_cast<uintptr_t>(new_vel) & 3u))
        return fail(LBM_ERR_ARG, "lbm_rebase_window: h and a 4-byte aligned new_vel are required");
    if (fmt != LBM_FMT_F16 && fmt != LBM_FMT_BF16) return fail(LBM_ERR_ARG, "lbm_rebase_window: bad format %d", fmt);
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_rebase_window: parity must be 0 or 1");
    {
        const char *a = reinterpret_cast<const char *>(P.wvel), *b = reinterpret_cast<const char *>(new_vel);
        const size_t n = sizeof(uint32_t) * (size_t)P.ww * (size_t)P.wh;
        if (a < b + n && b < a + n) return fail(LBM_ERR_ARG, "lbm_rebase_window: new_vel overlaps the current table");
    }
    const unsigned cells = (unsigned)P.ww * (unsigned)P.wh;
    const dim3 grid((cells + 255) / 256), block(256);
    if (fmt == LBM_FMT_F16)
        hipLaunchKernelGGL((k_rebase<FMT_F16R>), grid, block, 0, to_hip(stream), P, parity, h,
                           reinterpret_cast<uint32_t *>(new_vel));
    else
        hipLaunchKernelGGL((k_rebase<FMT_BF16R>), grid, block, 0, to_hip(stream), P, parity, h,
                           reinterpret_cast<uint32_t *>(new_vel));
    return check_launch("lbm_rebase_window");
}

int lbm_row_base(int H, const float *ux, const float *uy, float *table)
{
    if (H < 1 || !ux || !table) return fail(LBM_ERR_ARG, "lbm_row_base: H >= 1, ux and table are required");
    static const int CXh[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1}, CYh[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static const double Wh[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
    for (int y = 0; y < H; ++y) {
        const double u = ux[y], v = uy ? uy[y] : 0.0;
        const double usq = u * u + v * v;
        for (int q = 0; q < ROWB; ++q) {
            double d = 0.0;
            if (q < 9) {
                const double cu = CXh[q] * u + CYh[q] * v;
                d = Wh[q] * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
            }
            table[(size_t)ROWB * y + q] = (float)d;
        }
    }
    return LBM_OK;
}

size_t lbm_diagnostics_work_size(const lbm_params *) { return sizeof(double) * DIAG_BLOCKS * DIAG_N; }

int lbm_diagnostics(const lbm_params *p, const uint8_t *mask, const float *f, const float *inlet_u, int parity,
                    double *work, double *out, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0);
    if (rc) return rc;
    if (!mask || !f || !work || !out) return fail(LBM_ERR_ARG, "lbm_diagnostics: NULL pointer");
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_diagnostics: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_diagnostics: parity must be 0 or 1");
    hipStream_t s = to_hip(stream);
    if (p->bc == LBM_BC_PERIODIC)
        hipLaunchKernelGGL((k_diag_partial<true>), dim3(DIAG_BLOCKS), dim3(DIAG_THREADS), 0, s, P, mask, f, inlet_u, parity,
                           work);
    else
        hipLaunchKernelGGL((k_diag_partial<false>), dim3(DIAG_BLOCKS), dim3(DIAG_THREADS), 0, s, P, mask, f, inlet_u,
                           parity, work);
    rc = check_launch("lbm_diagnostics partial");
    if (rc) return rc;
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(DIAG_THREADS), 0, s, work, out);
    return check_launch("lbm_diagnostics final");
}

size_t lbm_diagnostics_scratch_bytes(void) { return SCRATCH_BYTES; }

int lbm_diagnostics_noalloc(const lbm_params *p, const uint8_t *mask, const float *f, const float *inlet_u,
                            int parity, double *out, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 0); // any lattice the library accepts: the partials do not grow with it
    if (rc) return rc;
    if (lbm_diagnostics_work_size(p) > SCRATCH_BYTES)
        return fail(LBM_ERR_SIZE, "lbm_diagnostics_noalloc: %zu bytes of partial sums exceed the %zu-byte scratch",
                    lbm_diagnostics_work_size(p), SCRATCH_BYTES);
    double *work = scratch_ptr(rc);
    if (!work) return rc;
    // every partial is overwritten by its workgroup in each call, so nothing needs clearing
    return lbm_diagnostics(p, mask, f, inlet_u, parity, work, out, stream);
}

} // extern "C"
