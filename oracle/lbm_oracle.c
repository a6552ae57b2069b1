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
} oracle_params;

/* ---------------------------- FP64 oracle ---------------------------- */
#define REAL double
#define SUFFIX _f64
#define FMA fma
#define SQRT sqrt
#define C_2_3 (2.0 / 3.0)
#define C_1_6 (1.0 / 6.0)
#define C_INV9 (1.0 / 9.0)
#define C_INV36 (1.0 / 36.0)
#define C_W0 (4.0 / 9.0)
#define C_W1 (1.0 / 9.0)
#define C_W5 (1.0 / 36.0)
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef FMA
#undef SQRT
#undef C_2_3
#undef C_1_6
#undef C_INV9
#undef C_INV36
#undef C_W0
#undef C_W1
#undef C_W5

/* ---------------------------- FP32 mirror ---------------------------- */
/* The literals are the reference's FP32 constants (lbm_core4.cu:15-16,55-57; lbm_core_smemaa.cu:4-7). */
#define REAL float
#define SUFFIX _f32
#define FMA fmaf
#define SQRT sqrtf
#define C_2_3 0.666666667f
#define C_1_6 0.16666666666f
#define C_INV9 0.1111111111f
#define C_INV36 0.02777777778f
#define C_W0 0.444444444f
#define C_W1 0.1111111111f
#define C_W5 0.0277777778f
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

int oracle_abi_version(void) { return 1; }
