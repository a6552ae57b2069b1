/*
 * D2Q9 oracle -- TEST INFRASTRUCTURE ONLY (see oracle_impl.h for what it restates and where).
 *
 * Builds oracle/liblbm_oracle.so with gcc -O2 -fopenmp -ffp-contract=off.  Exports, per precision
 * (_f64 = FP64 oracle, _f32 = FP32 mirror of the gfx950 operation order):
 *   oracle_aa_step_*    one in-place AA step (parity 0: local, parity 1: gather/scatter)
 *   oracle_pull_step_*  the two-buffer pull step of lbm_core4.cu:120-211
 *   oracle_macro_*      rho/u of a stored state (lbm_core_smemaa_endfused.cu:377-422)
 *   oracle_collide_*    the collision of one cell
 * plus FP64 helpers for the MRT transform and FP64 diagnostics (mass, momentum, momentum-exchange force).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

/* D2Q9 velocity set and opposite table, lbm_core4.cu:1-3 */
static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

enum { ORACLE_MRT = 0, ORACLE_BGK = 1 };
enum { ORACLE_CHANNEL = 0, ORACLE_PERIODIC = 1 };

/* same field layout as lbm_params in include/lbm.h (kept separate on purpose: the oracle stands alone) */
typedef struct {
    int32_t W, H;
    int32_t model;
    int32_t bc;
    int32_t clamp;
    float tau_inv;
    float cs;
    float s_e, s_eps, s_q;
    int32_t sums; /* 1: compensated (Neumaier) sums of rho - 1 and j in the MRT collision */
} oracle_params;

/* ---------------------------- FP64 oracle ---------------------------- */
#define REAL double
#define SUFFIX _f64
#define FMA fma
#define SQRT sqrt
#define FABS fabs
#define C_2_3 (2.0 / 3.0)
#define C_1_6 (1.0 / 6.0)
#define C_INV9 (1.0 / 9.0)
#define C_INV36 (1.0 / 36.0)
#define C_W0 (4.0 / 9.0)
#define C_W1 (1.0 / 9.0)
#define C_W5 (1.0 / 36.0)
#define C_1_12 (1.0 / 12.0)
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef FMA
#undef SQRT
#undef FABS
#undef C_2_3
#undef C_1_6
#undef C_INV9
#undef C_INV36
#undef C_W0
#undef C_W1
#undef C_W5
#undef C_1_12

/* ---------------------------- FP32 mirror ---------------------------- */
/* The literals are the reference's FP32 constants (lbm_core4.cu:15-16,55-57; lbm_core_smemaa.cu:4-7). */
#define REAL float
#define SUFFIX _f32
#define FMA fmaf
#define SQRT sqrtf
#define FABS fabsf
#define C_2_3 0.666666667f
#define C_1_6 0.16666666666f
#define C_INV9 0.1111111111f
#define C_INV36 0.02777777778f
#define C_W0 0.444444444f
#define C_W1 0.1111111111f
#define C_W5 0.0277777778f
#define C_1_12 0.0833333333f
#include "oracle_impl.h"

/* ---------------------------- FP64 helpers ---------------------------- */

/* MRT forward transform m = M f of lbm_core4.cu:152-165 with the momentum rows taken from f (jx, jy). */
void oracle_mrt_forward_f64(const double *f, double *m)
{
    double f14 = f[1] + f[2] + f[3] + f[4], f58 = f[5] + f[6] + f[7] + f[8];
    m[0] = f[0] + f14 + f58;
    m[1] = -4.0 * f[0] + 2.0 * f58 - f14;
    m[2] = 4.0 * f[0] + f58 - 2.0 * f14;
    m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    m[4] = -2.0 * (f[1] - f[3]) + f[5] - f[6] - f[7] + f[8];
    m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m[6] = -2.0 * (f[2] - f[4]) + f[5] + f[6] - f[7] - f[8];
    m[7] = f[1] - f[2] + f[3] - f[4];
    m[8] = f[5] - f[6] + f[7] - f[8];
}

/* MRT inverse f = M^-1 m of lbm_core4.cu:187-199 */
void oracle_mrt_inverse_f64(const double *m, double *f)
{
    double t0 = m[0] - m[1] + m[2];
    double t1 = 4.0 * m[0] - m[1] - 2.0 * m[2];
    double t2 = 4.0 * m[0] + 2.0 * m[1] + m[2];
    f[0] = t0 / 9.0;
    f[1] = (t1 + 6.0 * m[3] - 6.0 * m[4] + 9.0 * m[7]) / 36.0;
    f[2] = (t1 + 6.0 * m[5] - 6.0 * m[6] - 9.0 * m[7]) / 36.0;
    f[3] = (t1 - 6.0 * m[3] + 6.0 * m[4] + 9.0 * m[7]) / 36.0;
    f[4] = (t1 - 6.0 * m[5] + 6.0 * m[6] - 9.0 * m[7]) / 36.0;
    f[5] = (t2 + 6.0 * m[3] + 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] + 9.0 * m[8]) / 36.0;
    f[6] = (t2 - 6.0 * m[3] - 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] - 9.0 * m[8]) / 36.0;
    f[7] = (t2 - 6.0 * m[3] - 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] + 9.0 * m[8]) / 36.0;
    f[8] = (t2 + 6.0 * m[3] + 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] - 9.0 * m[8]) / 36.0;
}

/*
 * FP64 diagnostics of a float state (the GPU's storage type), summed in row order:
 *   out[0] = mass, out[1..2] = momentum over updated cells (same cell rule and Zou-He completion as
 *   oracle_macro), out[3..4] = momentum-exchange force on interior solid cells (rows 1..H-2):
 *   F = sum over links fluid x -> solid x+c_q of 2 c_q f*_q(x), where f*_q(x) sits in slot opp(q) at x in
 *   both parities (reversed: the post-collision value; canonical: the value bounced back by the wall).
 */
void oracle_diagnostics_f32(const oracle_params *pp, const uint8_t *mask, const float *f, const float *inlet_u,
                            int parity, double *out)
{
    kpar_t_f64 P;
    make_kpar_f64(pp, &P);
    const long plane = (long)P.W * P.H;
    double mass = 0, jx = 0, jy = 0, fx = 0, fy = 0;
    for (int y = 0; y < P.H; ++y) {
        for (int x = 0; x < P.W; ++x) {
            int r = role_f64(&P, mask, x, y);
            if (!r) continue;
            const long idx = (long)y * P.W + x;
            double g[9];
            for (int q = 0; q < 9; ++q) g[q] = (double)f[(parity ? OPP[q] : q) * plane + idx];
            if (r == 2 && parity == 0) {
                double u = (double)inlet_u[y];
                double rho = (g[0] + g[2] + g[4] + 2.0 * (g[3] + g[6] + g[7])) / (1.0 - u);
                mass += rho;
                jx += rho * u;
            } else {
                for (int q = 0; q < 9; ++q) {
                    mass += g[q];
                    jx += g[q] * CX[q];
                    jy += g[q] * CY[q];
                }
            }
            for (int q = 1; q < 9; ++q) {
                long d = nbr_f64(&P, x, y, CX[q], CY[q]);
                if (d < 0) continue;
                int dy = (int)(d / P.W);
                if (!mask[d]) continue;
                if (P.bc == ORACLE_CHANNEL && (dy == 0 || dy == P.H - 1)) continue;
                double fs = (double)f[OPP[q] * plane + idx];
                fx += 2.0 * CX[q] * fs;
                fy += 2.0 * CY[q] * fs;
            }
        }
    }
    out[0] = mass;
    out[1] = jx;
    out[2] = jy;
    out[3] = fx;
    out[4] = fy;
}

/* ---------------------------- 16-bit deviation storage ---------------------------- */
/* Memory holds h = f - w as IEEE half (fmt 1) or bfloat16 (fmt 2); the collision is the FP32 mirror on
 * deviations; stores clamp h to [-w, 10 - w] (the population clamp [0, 10], lbm_core4.cu:203) and round
 * to nearest even. */

static float half_to_float(uint16_t b)
{
    uint32_t sign = (uint32_t)(b >> 15) << 31, e = (b >> 10) & 0x1f, m = b & 0x3ff, u;
    if (e == 0) {
        if (m == 0) {
            u = sign;
        } else { /* subnormal: m * 2^-24 */
            float v = (float)m * 5.9604644775390625e-08f;
            memcpy(&u, &v, 4);
            u |= sign;
        }
    } else if (e == 31) {
        u = sign | 0x7f800000u | (m << 13);
    } else {
        u = sign | ((e + 112) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static uint16_t float_to_half(float f) /* round to nearest even */
{
    uint32_t u;
    memcpy(&u, &f, 4);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000);
    uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return sign | (a > 0x7f800000u ? 0x7e00 : 0x7c00);
    if (a >= 0x477ff000u) return sign | 0x7c00; /* rounds to >= 65520: infinity */
    if (a < 0x38800000u) {                      /* below 2^-14: subnormal half (or zero) */
        float af;
        memcpy(&af, &a, 4);
        float scaled = af * 16777216.0f;       /* exact: units of 2^-24 */
        float r = nearbyintf(scaled);          /* ties to even under the default rounding mode */
        return sign | (uint16_t)r;
    }
    uint32_t mant = a & 0x7fffff, exp = (a >> 23) - 112;
    uint32_t h = (exp << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fff;
    if (rem > 0x1000 || (rem == 0x1000 && (h & 1))) h += 1;
    return sign | (uint16_t)h;
}

static float bf16_to_float(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static uint16_t float_to_bf16(float f) /* round to nearest even */
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static inline float h16_load(int fmt, uint16_t b) { return fmt == 2 ? bf16_to_float(b) : half_to_float(b); }

static inline uint16_t h16_store(int fmt, int q, float h)
{
    float w = wq_f32(q);
    float lo = -w, hi = 10.0f - w;
    float c = clampr_f32(h, lo, hi);
    return fmt == 2 ? float_to_bf16(c) : float_to_half(c);
}

/* Per-row base of the 16-bit formats (the kernels' lbm_params.row_base): a slot on row y holding direction q stores
   the clamped deviation minus d = base[16 y + q] (one FP32 subtraction, then the 16-bit rounding) and loads the
   stored value plus d (one FP32 addition).  base == NULL: d = 0, the plain formats. */
#define ROWB 16
/* The base of a slot: a window of per-cell values (cells [x0, x0 + w) x [y0, y0 + h), ROWB floats per cell,
   row-major; base == NULL: none) takes precedence over the per-row table. */
typedef struct oracle_window {
    int x0, y0, w, h;
    const uint16_t *vel; /* (ux, uy) as IEEE halves per window cell, row-major */
} oracle_window;
static const oracle_window *g_win = 0; /* set for the duration of one oracle_aa_step_h16 call */

/* A window cell's base value of direction q: w_q ((3 cu + 4.5 cu^2) - 1.5 (ux^2 + uy^2)) in FP32 from its half
 * velocity, cu = c_q.u as one addition for the diagonals (the library's vel_base, lbm_device.h). */
static float vel_base_f32(int q, float ux, float uy)
{
    const float usq15 = 1.5f * (ux * ux + uy * uy);
    float cu;
    if (q == 0)
        cu = 0.0f;
    else if (CX[q] == 0)
        cu = CY[q] > 0 ? uy : -uy;
    else if (CY[q] == 0)
        cu = CX[q] > 0 ? ux : -ux;
    else
        cu = (CX[q] > 0 ? ux : -ux) + (CY[q] > 0 ? uy : -uy);
    return wq_f32(q) * ((3.0f * cu + 4.5f * (cu * cu)) - usq15);
}

/* base value of direction q for a slot of cell (x, y): the window cell's own, else its row's */
static inline float base_of(const float *rows, int q, int x, int y)
{
    const oracle_window *W = g_win;
    if (W && W->vel && x >= W->x0 && x < W->x0 + W->w && y >= W->y0 && y < W->y0 + W->h) {
        const uint16_t *v = W->vel + 2 * ((long)(y - W->y0) * W->w + (x - W->x0));
        return vel_base_f32(q, half_to_float(v[0]), half_to_float(v[1]));
    }
    return rows[ROWB * y + q];
}

static inline float h16_load_b(int fmt, uint16_t b, const float *base, int q, int x, int y)
{
    return base ? h16_load(fmt, b) + base_of(base, q, x, y) : h16_load(fmt, b);
}

static inline uint16_t h16_store_b(int fmt, int q, float h, const float *base, int x, int y)
{
    if (!base) return h16_store(fmt, q, h);
    float w = wq_f32(q);
    float c = clampr_f32(h, -w, 10.0f - w) - base_of(base, q, x, y);
    return fmt == 2 ? float_to_bf16(c) : float_to_half(c);
}

void oracle_convert_f32_h16(const float *x, uint16_t *y, long n, int fmt)
{
    for (long i = 0; i < n; ++i) y[i] = fmt == 2 ? float_to_bf16(x[i]) : float_to_half(x[i]);
}

void oracle_convert_h16_f32(const uint16_t *x, float *y, long n, int fmt)
{
    for (long i = 0; i < n; ++i) y[i] = h16_load(fmt, x[i]);
}

/* One AA step in place on 16-bit deviations; same streaming as oracle_aa_step_f32.  base: NULL, or the per-row
   base table (16 floats per row) that the row-based formats apply to every load and store. */
void oracle_aa_step_h16(const oracle_params *pp, const uint8_t *mask, uint16_t *f, int fmt, const float *inlet_u,
                        int parity, float *oux, float *ouy, float *orho, const float *base, const oracle_window *win)
{
    g_win = win;
    kpar_t_f32 P;
    make_kpar_f32(pp, &P);
    const long plane = (long)P.W * P.H;
    const int channel = P.bc == ORACLE_CHANNEL;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < P.H; ++y) {
        for (int x = 0; x < P.W; ++x) {
            int r = role_f32(&P, mask, x, y);
            if (!r) continue;
            const int inlet = r == 2;
            const long idx = (long)y * P.W + x;
            float g[9], mac[3];
            if (parity == 0) {
                for (int q = 0; q < 9; ++q) g[q] = h16_load_b(fmt, f[q * plane + idx], base, q, x, y);
            } else {
                for (int q = 0; q < 9; ++q) {
                    if (inlet && CX[q] == 1) {
                        g[q] = -wq_f32(q);
                        continue;
                    }
                    long s = nbr_f32(&P, x, y, -CX[q], -CY[q]);
                    /* bounce-back: this cell's slot q holds direction opp(q); else the source's slot opp(q), on the
                       source's row, holds direction q */
                    g[q] = mask[s] ? h16_load_b(fmt, f[q * plane + idx], base, OPP[q], x, y)
                                   : h16_load_b(fmt, f[OPP[q] * plane + s], base, q, (int)(s % P.W), (int)(s / P.W));
                }
            }
            if (P.model == ORACLE_BGK)
                collide_bgk_f32(&P, g, inlet, inlet ? inlet_u[y] : 0.0f, mac);
            else
                collide_mrt_f32(&P, g, inlet, inlet ? inlet_u[y] : 0.0f, mac, -1.0f);
            if (parity == 0) {
                for (int q = 0; q < 9; ++q) f[OPP[q] * plane + idx] = h16_store_b(fmt, q, g[q], base, x, y);
                if (channel && x == P.W - 2)
                    for (int q = 0; q < 9; ++q) f[OPP[q] * plane + idx + 1] = h16_store_b(fmt, q, g[q], base, x + 1, y);
            } else {
                for (int q = 0; q < 9; ++q) {
                    if (inlet && CX[q] == -1) continue;
                    long d = nbr_f32(&P, x, y, CX[q], CY[q]);
                    /* bounce-back: this cell's slot opp(q) then holds direction opp(q); else the target's slot q, on
                       the target's row (the clamp interval is the same for q and opp(q)) */
                    if (mask[d])
                        f[OPP[q] * plane + idx] = h16_store_b(fmt, OPP[q], g[q], base, x, y);
                    else
                        f[q * plane + d] = h16_store_b(fmt, q, g[q], base, (int)(d % P.W), (int)(d / P.W));
                }
                if (channel && x == P.W - 2) {
                    for (int q = 0; q < 9; ++q) {
                        if (CX[q] != -1) continue;
                        long d = (long)(y + CY[q]) * P.W + x;
                        if (!mask[d]) f[q * plane + d] = h16_store_b(fmt, q, g[q], base, x, y + CY[q]);
                    }
                }
            }
            if (orho) {
                orho[idx] = mac[0];
                oux[idx] = mac[1];
                ouy[idx] = mac[2];
            }
        }
    }
}

/* The kernels' base-window re-derivation (lbm_rebase_window), the same FP32 operations in the same order: per
   window cell decode with the current base, velocity u = j / rho (0 where rho is outside (0.5, 2) or u is not
   finite), new base d_q = w_q ((3 c.u + 4.5 (c.u)^2) - 1.5 u.u), re-encode as the conversion does:
   round16((f - w_q) - d). */
/* lbm_rebase_window's re-derivation on a 16-bit state in place: per window cell, decode with the current window
 * velocity, u = j / rho in FP32 (0 where rho is outside (0.5, 2) or |u_x| or |u_y| is not below 1), rounded to
 * half into new_vel, and the words re-encoded against the base values of the rounded velocity. */
void oracle_rebase_window(int W, int H, uint16_t *f, int fmt, int parity, const oracle_window *win, uint16_t *new_vel)
{
    const long plane = (long)W * H;
    (void)H;
    for (int j = 0; j < win->h; ++j)
        for (int i = 0; i < win->w; ++i) {
            const int x = win->x0 + i, y = win->y0 + j;
            const long idx = (long)y * W + x, cell = (long)j * win->w + i;
            const float ox = half_to_float(win->vel[2 * cell]), oy = half_to_float(win->vel[2 * cell + 1]);
            float fq[9];
            for (int q = 0; q < 9; ++q) {
                const int slot = parity ? OPP[q] : q;
                fq[q] = wq_f32(q) + (h16_load(fmt, f[slot * plane + idx]) + vel_base_f32(q, ox, oy));
            }
            const float rho = (((fq[0] + fq[1]) + (fq[2] + fq[3])) + ((fq[4] + fq[5]) + (fq[6] + fq[7]))) + fq[8];
            const float jx = ((fq[1] - fq[3]) + (fq[5] - fq[6])) + (fq[8] - fq[7]);
            const float jy = ((fq[2] - fq[4]) + (fq[5] + fq[6])) - (fq[7] + fq[8]);
            const float ux = jx / rho, uy = jy / rho;
            const int ok = rho > 0.5f && rho < 2.0f && fabsf(ux) < 1.0f && fabsf(uy) < 1.0f;
            const uint16_t hx = ok ? float_to_half(ux) : 0, hy = ok ? float_to_half(uy) : 0;
            new_vel[2 * cell] = hx;
            new_vel[2 * cell + 1] = hy;
            const float nx = half_to_float(hx), ny = half_to_float(hy);
            for (int q = 0; q < 9; ++q) {
                const float v = (fq[q] - wq_f32(q)) - vel_base_f32(q, nx, ny);
                const int slot = parity ? OPP[q] : q;
                f[slot * plane + idx] = fmt == 2 ? float_to_bf16(v) : float_to_half(v);
            }
        }
}

int oracle_abi_version(void) { return 1; }

/* Threads of the OpenMP loops (<= 0: the runtime's default); returns the count now in effect. */
#include <omp.h>
int oracle_set_threads(int n)
{
    static int dflt = 0;
    if (!dflt) dflt = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : dflt);
    return omp_get_max_threads();
}


/* Neumaier-compensated FP32 sum of x[0..n) in index order (the summation the collision uses with sums = 1);
   exported for the tests that pin it against an exactly rounded sum. */
float oracle_neumaier_f32(const float *x, int n)
{
    if (n <= 0) return 0.0f;
    float s = x[0], c = 0.0f;
    for (int i = 1; i < n; ++i) neu_add_f32(&s, &c, x[i]);
    return s + c;
}
