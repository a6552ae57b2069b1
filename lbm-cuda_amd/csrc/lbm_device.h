// Device-side D2Q9 collision shared by every gfx950 kernel of liblbm.
//
// One __device__ collide function is used by the single-step, pull and temporally blocked kernels, so
// that k fused steps are bit-identical to k single steps.  The code is compiled with -ffp-contract=off:
// every fused multiply-add is written out (__builtin_fmaf), divisions are the correctly rounded __fdiv_rn
// and square roots sqrtf (which hipcc expands to a correctly rounded sequence on gfx950; __fsqrt_rn lowers
// to the bare v_sqrt_f32, which is not), and nothing is reassociated.
//
// Physics follows lbm_core4.cu:29-210 (MRT + Smagorinsky, Zou-He inlet, clamps) and
// lbm_core_smemaa.cu:159-220 (BGK); the operation order is this project's own (fewer VALU operations:
// shared partial sums, constants folded into FMAs).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// D2Q9 velocity set and opposite table (lbm_core4.cu:1-3)
__host__ __device__ constexpr int cx(int q) { return q == 1 || q == 5 || q == 8 ? 1 : (q == 3 || q == 6 || q == 7 ? -1 : 0); }
__host__ __device__ constexpr int cy(int q) { return q == 2 || q == 5 || q == 6 ? 1 : (q == 4 || q == 7 || q == 8 ? -1 : 0); }
__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : (q <= 4 ? ((q + 1) & 3) + 1 : ((q - 3) & 3) + 5); }
static_assert(opp(1) == 3 && opp(2) == 4 && opp(3) == 1 && opp(4) == 2, "opp");
static_assert(opp(5) == 7 && opp(6) == 8 && opp(7) == 5 && opp(8) == 6, "opp");

enum { MODEL_MRT = 0, MODEL_BGK = 1 };

// Kernel-side parameters, derived on the host in FP32 (tau0 = 1/tau_inv etc., as lbm_core4.cu:175-176).
struct KPar {
    int W, H;
    int plane; // W * H
    int les;   // MRT: Smagorinsky on (cs != 0)
    int clamp; // MRT: velocity and density clamps
    float tau_inv, tau0, tau0sq, c18cs2, s_e, s_eps, s_q;
};

// The reference's FP32 literals (lbm_core4.cu:15-16,55-57; lbm_core_smemaa.cu:4-7)
#define LBM_C_2_3 0.666666667f
#define LBM_C_1_6 0.16666666666f
#define LBM_INV9 0.1111111111f
#define LBM_INV36 0.02777777778f
#define LBM_W0 0.444444444f
#define LBM_W1 0.1111111111f
#define LBM_W5 0.0277777778f

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// Zou-He velocity inlet (lbm_core4.cu:54-57): rebuilds f1, f5, f8 from the known populations.
__device__ __forceinline__ float zou_he(float f[9], float u)
{
    float num = fma_(2.0f, (f[3] + f[6]) + f[7], (f[0] + f[2]) + f[4]);
    float rho = __fdiv_rn(num, 1.0f - u);
    float ru = rho * u;
    float h = 0.5f * (f[2] - f[4]);
    f[1] = fma_(LBM_C_2_3, ru, f[3]);
    f[5] = fma_(LBM_C_1_6, ru, f[7] - h);
    f[8] = fma_(LBM_C_1_6, ru, f[6] + h);
    return rho;
}

// MRT + Smagorinsky collision (lbm_core4.cu:124-210).  `inlet` selects Zou-He and the inlet collision
// of lbm_core4.cu:54-117 (no LES, no clamps, m3 = rho u, m5 = 0).  mac = (rho, ux, uy) as written by the
// reference.  f is replaced by the clamped post-collision populations.
__device__ __forceinline__ void collide_mrt(const KPar &P, float f[9], bool inlet, float uin, float mac[3])
{
    float rho = 0.f, ux = 0.f, uy = 0.f, rc = 0.f;
    bool les = false;
    if (inlet) {
        rho = zou_he(f, uin);
        ux = uin;
        uy = 0.0f;
        rc = rho;
    }
    const float s13 = f[1] + f[3], s24 = f[2] + f[4], s57 = f[5] + f[7], s68 = f[6] + f[8];
    const float f14 = s13 + s24, f58 = s57 + s68;
    const float a13 = f[1] - f[3], a24 = f[2] - f[4], a57 = f[5] - f[7], a86 = f[8] - f[6];
    const float sx = a57 + a86; // f5 - f6 - f7 + f8
    const float sy = a57 - a86; // f5 + f6 - f7 - f8
    if (!inlet) {
        rho = (f[0] + f14) + f58;
        const float jx = a13 + sx, jy = a24 + sy;
        const float inv_rho = __fdiv_rn(1.0f, rho);
        ux = jx * inv_rho;
        uy = jy * inv_rho;
        rc = rho;
        if (P.clamp) {
            ux = clampf(ux, -0.4f, 0.4f);
            uy = clampf(uy, -0.4f, 0.4f);
            rc = clampf(rho, 0.0f, 1.5f);
        }
        les = P.les != 0;
    }
    mac[0] = rc;
    mac[1] = ux;
    mac[2] = uy;
    // forward transform, Lallemand-Luo order (lbm_core4.cu:152-165)
    const float m0 = rc;
    float m1 = fma_(2.0f, f58, fma_(-4.0f, f[0], -f14));
    float m2 = fma_(-2.0f, f14, fma_(4.0f, f[0], f58));
    const float m3 = rc * ux;
    float m4 = fma_(-2.0f, a13, sx);
    const float m5 = rc * uy;
    float m6 = fma_(-2.0f, a24, sy);
    float m7 = s13 - s24;
    float m8 = s57 - s68;
    // equilibria (lbm_core4.cu:166-170)
    const float ux2 = ux * ux, uy2 = uy * uy;
    const float usq = ux2 + uy2;
    const float m1eq = fma_(3.0f, usq, -2.0f) * rc;
    const float m2eq = fma_(-3.0f, usq, 1.0f) * rc;
    const float m7eq = rc * (ux2 - uy2);
    const float m8eq = m3 * uy;
    const float n7 = m7 - m7eq;
    const float n8 = m8 - m8eq;
    float w = P.tau_inv;
    if (les) {
        // Smagorinsky (lbm_core4.cu:172-177)
        const float Q = sqrtf(fma_(n7, n7, fma_(n8, n8, 1e-8f)));
        const float inv_rc = __fdiv_rn(1.0f, rc);
        const float d = fma_(P.c18cs2, Q * inv_rc, P.tau0sq);
        const float tau_eff = 0.5f * (P.tau0 + sqrtf(d));
        w = __fdiv_rn(1.0f, tau_eff);
    }
    // relaxation (lbm_core4.cu:180-185)
    m1 = fma_(-P.s_e, m1 - m1eq, m1);
    m2 = fma_(-P.s_eps, m2 - m2eq, m2);
    m4 = fma_(-P.s_q, m4 + m3, m4);
    m6 = fma_(-P.s_q, m6 + m5, m6);
    m7 = fma_(-w, n7, m7);
    m8 = fma_(-w, n8, m8);
    // inverse (lbm_core4.cu:187-199)
    const float t0 = (m0 - m1) + m2;
    const float t1 = fma_(-2.0f, m2, fma_(4.0f, m0, -m1));
    const float t2 = fma_(4.0f, m0, fma_(2.0f, m1, m2));
    const float A = (m3 - m4) * 6.0f;
    const float B = (m5 - m6) * 6.0f;
    const float Pp = m7 * 9.0f;
    const float X = fma_(6.0f, m3, m4 * 3.0f);
    const float Y = fma_(6.0f, m5, m6 * 3.0f);
    const float Z = m8 * 9.0f;
    const float t1p = t1 + Pp, t1m = t1 - Pp;
    const float u = t2 + Z, v = t2 - Z;
    const float XpY = X + Y, XmY = X - Y;
    f[0] = clampf(t0 * LBM_INV9, 0.0f, 10.0f);
    f[1] = clampf((t1p + A) * LBM_INV36, 0.0f, 10.0f);
    f[2] = clampf((t1m + B) * LBM_INV36, 0.0f, 10.0f);
    f[3] = clampf((t1p - A) * LBM_INV36, 0.0f, 10.0f);
    f[4] = clampf((t1m - B) * LBM_INV36, 0.0f, 10.0f);
    f[5] = clampf((u + XpY) * LBM_INV36, 0.0f, 10.0f);
    f[6] = clampf((v - XmY) * LBM_INV36, 0.0f, 10.0f);
    f[7] = clampf((u - XpY) * LBM_INV36, 0.0f, 10.0f);
    f[8] = clampf((v + XmY) * LBM_INV36, 0.0f, 10.0f);
}

// BGK collision (lbm_core_smemaa.cu:159-220): no velocity/density clamps, populations clamped to [0,10].
__device__ __forceinline__ void collide_bgk(const KPar &P, float f[9], bool inlet, float uin, float mac[3])
{
    float rho, ux, uy;
    if (inlet) {
        rho = zou_he(f, uin);
        ux = uin;
        uy = 0.0f;
    } else {
        const float s13 = f[1] + f[3], s24 = f[2] + f[4], s57 = f[5] + f[7], s68 = f[6] + f[8];
        const float a13 = f[1] - f[3], a24 = f[2] - f[4], a57 = f[5] - f[7], a86 = f[8] - f[6];
        rho = (f[0] + (s13 + s24)) + (s57 + s68);
        const float inv_rho = __fdiv_rn(1.0f, rho);
        ux = (a13 + (a57 + a86)) * inv_rho;
        uy = (a24 + (a57 - a86)) * inv_rho;
    }
    mac[0] = rho;
    mac[1] = ux;
    mac[2] = uy;
    const float usq = fma_(uy, uy, ux * ux);
    const float base = fma_(-1.5f, usq, 1.0f);
    const float wr0 = LBM_W0 * rho, wr1 = LBM_W1 * rho, wr5 = LBM_W5 * rho;
    const float cu[9] = {0.0f, ux, uy, -ux, -uy, ux + uy, uy - ux, -(ux + uy), ux - uy};
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float wr = q == 0 ? wr0 : (q < 5 ? wr1 : wr5);
        const float poly = q == 0 ? base : fma_(cu[q], fma_(4.5f, cu[q], 3.0f), base);
        const float feq = wr * poly;
        f[q] = clampf(fma_(P.tau_inv, feq - f[q], f[q]), 0.0f, 10.0f);
    }
}

template <int MODEL>
__device__ __forceinline__ void collide(const KPar &P, float f[9], bool inlet, float uin, float mac[3])
{
    if constexpr (MODEL == MODEL_BGK)
        collide_bgk(P, f, inlet, uin, mac);
    else
        collide_mrt(P, f, inlet, uin, mac);
}

} // namespace lbm
