// Device-side D2Q9 collision shared by every gfx950 kernel of liblbm.
//
// One __device__ collide function is used by the single-step, pull and temporally blocked kernels, so
// that k fused steps are bit-identical to k single steps.  The code is compiled with -ffp-contract=off:
// every fused multiply-add is written out (__builtin_fmaf), divisions are the correctly rounded __fdiv_rn
// and square roots are correctly rounded (sqrt_rn_normal; note that __fsqrt_rn lowers to the bare
// v_sqrt_f32, which is not), and nothing is reassociated.
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

// MODEL_MRT_SUMS: MRT + LES with compensated (Neumaier) sums of rho - 1 and j (lbm_params.sums = 1).
// MODEL_MRT_CLAMP0 / _CLAMP1: MODEL_MRT with the velocity/density clamps known off / on at compile time (the
// same arithmetic as MODEL_MRT with that lbm_params.clamp, without the run-time select).
// MODEL_FAST (a flag or-ed into any of them but MODEL_MRT_SUMS): the collision's reciprocals and square roots are the
// hardware's v_rcp_f32 / v_sqrt_f32 (within 1 ulp; lbm_params.fast), as the reference's -use_fast_math build computes
// them, instead of the correctly rounded forms the CPU mirror reproduces bit for bit.
enum { MODEL_MRT = 0, MODEL_BGK = 1, MODEL_MRT_SUMS = 2, MODEL_MRT_CLAMP0 = 3, MODEL_MRT_CLAMP1 = 4, MODEL_FAST = 8 };
__host__ __device__ constexpr int model_base(int m) { return m & ~MODEL_FAST; }
__host__ __device__ constexpr bool model_fast(int m) { return (m & MODEL_FAST) != 0; }

// Kernel-side parameters, derived on the host in FP32 (tau0 = 1/tau_inv etc., as lbm_core4.cu:175-176).
struct KPar {
    int W, H;
    int plane; // W * H
    int les;   // MRT: Smagorinsky on (cs != 0)
    int clamp; // MRT: velocity and density clamps
    float tau_inv, tau0, tau0sq, c18cs2, s_e, s_eps, s_q;
    int sums; // lbm_params.sums: compensated sums (MODEL_MRT_SUMS instantiations)
    const float *base; // lbm_params.row_base: per-row base of the row-based 16-bit formats (ROWB floats per row)
    // lbm_params.window: cells [wx0, wx0 + ww) x [wy0, wy0 + wh) take their base values from the equilibrium of a
    // velocity of their own, wvel (ux, uy as two IEEE halves per cell, ux in the low half, row-major over the
    // window), instead of from their row; ww = 0: no window
    int wx0, wy0, ww, wh;
    const uint32_t *wvel;
    // two-step kernel: the tiles meeting the window, [wtx0, wtx0 + wtw) x [wty0, wty0 + wth) in tile coordinates,
    // are dispatched first (wtw = 0: none)
    int wtx0, wty0, wtw, wth;
    int rowmap; // single step: the state (9 W H population bytes) exceeds the 256 MB MALL -> xcd_row_tile order
    int fast;   // lbm_params.fast: MODEL_FAST instantiations
};

// The reference's FP32 literals (lbm_core4.cu:15-16,55-57; lbm_core_smemaa.cu:4-7)
#define LBM_C_2_3 0.666666667f
#define LBM_C_1_6 0.16666666666f
#define LBM_INV9 0.1111111111f
#define LBM_INV36 0.02777777778f
#define LBM_W0 0.444444444f
#define LBM_W1 0.1111111111f
#define LBM_W5 0.0277777778f
#define LBM_C_1_12 0.0833333333f

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Correctly rounded square root for positive normal x (here x >= 1e-8): the hardware estimate (within 1 ulp)
// is moved to the neighbour whose residual x - s*s has the right sign -- the same correction hipcc emits for
// sqrtf, without the rescaling and special-value paths that only denormal, zero, infinite or NaN inputs need.
__device__ __forceinline__ float sqrt_rn_normal(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sdn = __builtin_bit_cast(float, __builtin_bit_cast(int, s) - 1);
    const float sup = __builtin_bit_cast(float, __builtin_bit_cast(int, s) + 1);
    const float rdn = fma_(-sdn, s, x);
    const float rup = fma_(-sup, s, x);
    const float t = rdn <= 0.0f ? sdn : s;
    return rup > 0.0f ? sup : t;
}
// Correctly rounded 1/x, equal to __fdiv_rn(1, x) for every x.  For 2^-125 <= |x| <= 2^125 the hardware
// estimate y (within 1 ulp) takes one Newton step with an exact residual, r = fma(-x, y, 1) and
// y + r y = fma(r, y, y) (Markstein's reciprocal iteration): 3 instructions instead of the 11 of the
// general division, which handles scaling and special values.  Other x (zero, denormal, huge, inf, NaN)
// take the general division.  lbm_selftest compares both over all 2^32 inputs.
__device__ __forceinline__ float recip_rn(float x)
{
    const float ax = __builtin_fabsf(x);
    if (ax >= 0x1p-125f && ax <= 0x1p125f) { // false for NaN
        const float y = __builtin_amdgcn_rcpf(x);
        return fma_(fma_(-x, y, 1.0f), y, y);
    }
    return __fdiv_rn(1.0f, x);
}
// FAST: the hardware estimates themselves (within 1 ulp); otherwise the correctly rounded forms above
template <bool FAST> __device__ __forceinline__ float recip_m(float x) { return FAST ? __builtin_amdgcn_rcpf(x) : recip_rn(x); }
template <bool FAST> __device__ __forceinline__ float sqrt_m(float x)
{
    return FAST ? __builtin_amdgcn_sqrtf(x) : sqrt_rn_normal(x);
}

// clamp to [lo, hi] (lo <= hi) in one v_med3_f32; equal to min(max(v, lo), hi) for every non-NaN v
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return __builtin_amdgcn_fmed3f(v, lo, hi); }

// The collision works on deviations h_q = f_q - w_q from the rest state (rho = 1, u = 0).  The rest
// state's moments (1, -2, 1, 0, ..., 0) are carried exactly and only the small remainders are rounded,
// which keeps the FP32 result within a few 1e-8 of FP64 and is the form FP16/BF16 storage keeps in memory.
// Every formula below is the reference's (cited), rewritten for deviations.
__device__ __forceinline__ float wq(int q) { return q == 0 ? LBM_W0 : (q < 5 ? LBM_W1 : LBM_W5); }

// Zou-He velocity inlet (lbm_core4.cu:54-57) on deviations: rebuilds h1, h5, h8; returns rho - 1.
// f0+f2+f4+2(f3+f6+f7) = 1 + (h0+h2+h4+2(h3+h6+h7)), so rho - 1 = (H + u) / (1 - u).
__device__ __forceinline__ float zou_he(float h[9], float u, float &rho)
{
    const float Hs = fma_(2.0f, (h[3] + h[6]) + h[7], (h[0] + h[2]) + h[4]);
    const float drho = __fdiv_rn(Hs + u, 1.0f - u);
    rho = 1.0f + drho;
    const float ru = rho * u;
    const float hh = 0.5f * (h[2] - h[4]);
    h[1] = fma_(LBM_C_2_3, ru, h[3]);
    h[5] = fma_(LBM_C_1_6, ru, h[7] - hh);
    h[8] = fma_(LBM_C_1_6, ru, h[6] + hh);
    return drho;
}

// Non-inline constants of the collision and of the 16-bit store clamp, and the two collision parameters that
// meet another scalar in one instruction.  A VOP3 instruction reads at most one scalar register or literal, so
// med3 / select / fma with two such operands needs one of them in a VGPR; in straight-line code the compiler
// rematerialises it (a v_mov) at every use.  Kernels take collide_consts(P); a kernel that collides many cells
// per thread in straight-line code takes opaque_consts(P) once instead: its values stay in VGPRs.
struct CollideK {
    float ulim = 0.4f;                                    // velocity clamp
    float inv15 = __builtin_bit_cast(float, 0x3F2AAAABu); // 1 / 1.5 rounded
    float inf = __builtin_inff();                         // 1 / 0
    float hi[3] = {10.0f - LBM_W0, 10.0f - LBM_W1, 10.0f - LBM_W5}; // the population clamp 10 - w_q per weight
    float tau0sq = 0.0f, tau_inv = 0.0f;                  // KPar::tau0sq, KPar::tau_inv
};
__device__ __forceinline__ CollideK collide_consts(const KPar &P)
{
    CollideK k;
    k.tau0sq = P.tau0sq;
    k.tau_inv = P.tau_inv;
    return k;
}
__device__ __forceinline__ CollideK opaque_consts(const KPar &P)
{
    CollideK k = collide_consts(P);
    asm volatile("" : "+v"(k.ulim), "+v"(k.inv15), "+v"(k.inf), "+v"(k.hi[0]), "+v"(k.hi[1]), "+v"(k.hi[2]));
    asm volatile("" : "+v"(k.tau0sq), "+v"(k.tau_inv));
    return k;
}
__device__ __forceinline__ constexpr int wclass(int q) { return q == 0 ? 0 : (q < 5 ? 1 : 2); }

// Neumaier's compensated summation (Neumaier 1974, the improved Kahan-Babuska scheme): s + t rounded, the
// rounding error of that addition accumulated in c; the sum is s + c at the end.
__device__ __forceinline__ void neu_add(float &s, float &c, float t)
{
    const float u = s + t;
    c += __builtin_fabsf(s) >= __builtin_fabsf(t) ? (s - u) + t : (t - u) + s;
    s = u;
}

// rho - 1, jx, jy as compensated sums over the deviations, terms in index order
__device__ __forceinline__ void moments_neu(const float h[9], float &drho, float &jx, float &jy)
{
    float s = h[0], c = 0.0f;
#pragma unroll
    for (int q = 1; q < 9; ++q) neu_add(s, c, h[q]);
    drho = s + c;
    s = h[1], c = 0.0f;
    neu_add(s, c, -h[3]);
    neu_add(s, c, h[5]);
    neu_add(s, c, -h[6]);
    neu_add(s, c, -h[7]);
    neu_add(s, c, h[8]);
    jx = s + c;
    s = h[2], c = 0.0f;
    neu_add(s, c, -h[4]);
    neu_add(s, c, h[5]);
    neu_add(s, c, h[6]);
    neu_add(s, c, -h[7]);
    neu_add(s, c, -h[8]);
    jy = s + c;
}

// MRT + Smagorinsky collision (lbm_core4.cu:124-199) on deviations, h -> h' (unclamped).  `inlet`
// selects Zou-He and the inlet collision of lbm_core4.cu:54-93 (no LES, no clamps, m3 = rho u, m5 = 0).
// mac = (rho, ux, uy) as written by the reference (lbm_core4.cu:59-61,145-147).
// EXT: the stress moments relax at w_ext when w_ext >= 0 (a learned closure's rate), else as usual.
// SUMS: rho - 1 and j as compensated sums (moments_neu) instead of the pairwise FP32 sums.
template <bool EXT = false, bool SUMS = false, int CLAMP = -1, bool FAST = false>
__device__ __forceinline__ void collide_mrt(const KPar &P, float h[9], bool inlet, float uin, float mac[3],
                                            float w_ext, const CollideK &K)
{
    // one if/else that defines every value on both sides: defaults set before the branch would be
    // materialised for every lane, while the inlet side runs only in waves that hold inlet cells
    float rho, ux, uy, drc, rc, inv_rc;
    float s13, s24, s57, s68, a13, a24, a57, a86, h14, h58, sx, sy;
    bool les;
    auto sums = [&]() {
        s13 = h[1] + h[3];
        s24 = h[2] + h[4];
        s57 = h[5] + h[7];
        s68 = h[6] + h[8];
        a13 = h[1] - h[3];
        a24 = h[2] - h[4];
        a57 = h[5] - h[7];
        a86 = h[8] - h[6];
        h14 = s13 + s24;
        h58 = s57 + s68;
        sx = a57 + a86; // f5 - f6 - f7 + f8
        sy = a57 - a86; // f5 + f6 - f7 - f8
    };
    if (inlet) {
        drc = zou_he(h, uin, rho);
        rc = rho;
        ux = uin;
        uy = 0.0f;
        inv_rc = 0.0f; // unused: no LES at the inlet
        les = false;
        sums();
    } else {
        sums();
        // moments and clamps (lbm_core4.cu:130-144); rho = 1 + drho
        float drho = (h[0] + h14) + h58;
        float jx = a13 + sx, jy = a24 + sy;
        if constexpr (SUMS) moments_neu(h, drho, jx, jy);
        rho = 1.0f + drho;
        const float inv_rho = recip_m<FAST>(rho);
        ux = jx * inv_rho;
        uy = jy * inv_rho;
        drc = drho;
        if (CLAMP > 0 || (CLAMP < 0 && P.clamp)) { // CLAMP: 1 on, 0 off, -1 lbm_params.clamp
            ux = clampf(ux, -K.ulim, K.ulim);
            uy = clampf(uy, -K.ulim, K.ulim);
            drc = clampf(drho, -1.0f, 0.5f); // rho in [0, 1.5]
        }
        rc = 1.0f + drc;
        // 1/rc == 1/rho bit for bit unless the density clamp engaged, and then rc is 1.5 or 0 exactly:
        // 1/1.5 rounds to 0x3F2AAAAB and 1/0 is +inf, so no second division is needed
        inv_rc = inv_rho;
        if (__builtin_expect(drc != drho, 0)) inv_rc = drc > 0.0f ? K.inv15 : K.inf; // the density clamp engaged
        les = P.les != 0;
    }
    mac[0] = rc;
    mac[1] = ux;
    mac[2] = uy;
    // forward transform (lbm_core4.cu:152-165): m1 = -2 + m1d, m2 = 1 + m2d, the rest carry no rest part
    float m1d = fma_(2.0f, h58, fma_(-4.0f, h[0], -h14));
    float m2d = fma_(-2.0f, h14, fma_(4.0f, h[0], h58));
    const float m3 = rc * ux;
    float m4 = fma_(-2.0f, a13, sx);
    const float m5 = rc * uy;
    float m6 = fma_(-2.0f, a24, sy);
    float m7 = s13 - s24;
    float m8 = s57 - s68;
    // non-equilibrium parts (lbm_core4.cu:166-170): m1 - (3u^2 - 2) rho = m1d + 2 (rho - 1) - 3u^2 rho,
    // m2 - (1 - 3u^2) rho = m2d - (rho - 1) + 3u^2 rho
    const float ux2 = ux * ux, uy2 = uy * uy;
    const float usq = ux2 + uy2;
    const float u3 = 3.0f * usq;
    const float n1 = fma_(-u3, rc, fma_(2.0f, drc, m1d));
    const float n2 = fma_(u3, rc, m2d - drc);
    const float m7eq = rc * (ux2 - uy2);
    const float m8eq = m3 * uy;
    const float n7 = m7 - m7eq;
    const float n8 = m8 - m8eq;
    float w = K.tau_inv;
    bool smag = les;
    if constexpr (EXT) {
        if (w_ext >= 0.0f) {
            w = w_ext;
            smag = false;
        }
    }
    if (smag) {
        // Smagorinsky (lbm_core4.cu:172-177)
        const float Q = sqrt_m<FAST>(fma_(n7, n7, fma_(n8, n8, 1e-8f)));
        const float d = fma_(P.c18cs2, Q * inv_rc, K.tau0sq);
        const float tau_eff = 0.5f * (P.tau0 + sqrt_m<FAST>(d));
        w = recip_m<FAST>(tau_eff);
    }
    // relaxation (lbm_core4.cu:180-185)
    m1d = fma_(-P.s_e, n1, m1d);
    m2d = fma_(-P.s_eps, n2, m2d);
    m4 = fma_(-P.s_q, m4 + m3, m4);
    m6 = fma_(-P.s_q, m6 + m5, m6);
    m7 = fma_(-w, n7, m7);
    m8 = fma_(-w, n8, m8);
    // inverse (lbm_core4.cu:187-199) of the deviation moments (rho - 1, m1d, m2d, m3, ..., m8), with the
    // common 1/36 folded into the coefficients: 6/36 = 1/6, 3/36 = 1/12, 9/36 = 1/4
    const float t0 = (drc - m1d) + m2d;
    const float t1 = fma_(-2.0f, m2d, fma_(4.0f, drc, -m1d));
    const float t2 = fma_(4.0f, drc, fma_(2.0f, m1d, m2d));
    const float q7 = m7 * 0.25f;
    const float q8 = m8 * 0.25f;
    const float t1p = fma_(t1, LBM_INV36, q7), t1m = fma_(t1, LBM_INV36, -q7);
    const float d34 = m3 - m4, d56 = m5 - m6;
    const float X = fma_(LBM_C_1_6, m3, m4 * LBM_C_1_12);
    const float Y = fma_(LBM_C_1_6, m5, m6 * LBM_C_1_12);
    const float u = fma_(t2, LBM_INV36, q8), v = fma_(t2, LBM_INV36, -q8);
    const float XpY = X + Y, XmY = X - Y;
    h[0] = t0 * LBM_INV9;
    h[1] = fma_(d34, LBM_C_1_6, t1p);
    h[2] = fma_(d56, LBM_C_1_6, t1m);
    h[3] = fma_(d34, -LBM_C_1_6, t1p);
    h[4] = fma_(d56, -LBM_C_1_6, t1m);
    h[5] = u + XpY;
    h[6] = v - XmY;
    h[7] = u - XpY;
    h[8] = v + XmY;
}

// BGK collision (lbm_core_smemaa.cu:159-212) on deviations: h' = h + tau_inv (heq - h) with
// heq_q = feq_q - w_q = w_q (drho + rho (3cu + 4.5cu^2 - 1.5u^2)).  No velocity/density clamps.
template <bool FAST = false>
__device__ __forceinline__ void collide_bgk(const KPar &P, float h[9], bool inlet, float uin, float mac[3])
{
    float rho, drho, ux, uy;
    if (inlet) {
        drho = zou_he(h, uin, rho);
        ux = uin;
        uy = 0.0f;
    } else {
        const float s13 = h[1] + h[3], s24 = h[2] + h[4], s57 = h[5] + h[7], s68 = h[6] + h[8];
        const float a13 = h[1] - h[3], a24 = h[2] - h[4], a57 = h[5] - h[7], a86 = h[8] - h[6];
        drho = (h[0] + (s13 + s24)) + (s57 + s68);
        rho = 1.0f + drho;
        const float inv_rho = recip_m<FAST>(rho);
        ux = (a13 + (a57 + a86)) * inv_rho;
        uy = (a24 + (a57 - a86)) * inv_rho;
    }
    mac[0] = rho;
    mac[1] = ux;
    mac[2] = uy;
    const float usq15 = 1.5f * fma_(uy, uy, ux * ux);
    const float cu[9] = {0.0f, ux, uy, -ux, -uy, ux + uy, uy - ux, -(ux + uy), ux - uy};
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float poly = q == 0 ? -usq15 : fma_(cu[q], fma_(4.5f, cu[q], 3.0f), -usq15);
        const float heq = wq(q) * fma_(rho, poly, drho);
        h[q] = fma_(P.tau_inv, heq - h[q], h[q]);
    }
}

template <int MODEL, bool EXT = false>
__device__ __forceinline__ void collide_dev(const KPar &P, float h[9], bool inlet, float uin, float mac[3],
                                            float w_ext, const CollideK &K)
{
    constexpr int M = model_base(MODEL);
    constexpr bool FAST = model_fast(MODEL);
    static_assert(!(FAST && M == MODEL_MRT_SUMS), "compensated sums come with correctly rounded arithmetic");
    if constexpr (M == MODEL_BGK)
        collide_bgk<FAST>(P, h, inlet, uin, mac);
    else
        collide_mrt<EXT, M == MODEL_MRT_SUMS, M == MODEL_MRT_CLAMP1 ? 1 : (M == MODEL_MRT_CLAMP0 ? 0 : -1), FAST>(
            P, h, inlet, uin, mac, w_ext, K);
}

template <int MODEL, bool EXT = false>
__device__ __forceinline__ void collide_dev(const KPar &P, float h[9], bool inlet, float uin, float mac[3],
                                            float w_ext = -1.0f)
{
    collide_dev<MODEL, EXT>(P, h, inlet, uin, mac, w_ext, collide_consts(P));
}

// ------------------------------------------------------------------------------------------------------
// Storage formats.  The kernels move deviations h = f - w between memory and the collision:
//   FMT_F32   memory holds f (FP32):              h = f - w,            f' = clamp(w + h', 0, 10)
//   FMT_F16   memory holds h as IEEE half:        h = (float)half,      half(clamp(h', -w, 10 - w))
//   FMT_BF16  memory holds h as bfloat16:         h = bf16 << 16,       bf16(clamp(h', -w, 10 - w))
// Both 16-bit conversions round to nearest even (v_cvt_f16_f32, v_cvt_pk_bf16_f32).  The population clamp
// [0, 10] of lbm_core4.cu:203 becomes the same interval shifted by -w.
// Memory is addressed through buffer resources: a scalar base, the plane offset as scalar offset and one
// 32-bit vector byte offset per access (no 64-bit vector address arithmetic); accesses past the end of
// the resource read 0 and write nothing.
// ------------------------------------------------------------------------------------------------------
enum { FMT_F32 = 0, FMT_F16 = 1, FMT_BF16 = 2, FMT_F16R = 3, FMT_BF16R = 4 };
// FMT_F16R / FMT_BF16R: the 16-bit formats relative to a per-row base (lbm_params.row_base, KPar::base).  A slot on
// row y that holds a population moving in direction q stores half(clamp(h) - d_q(y)) instead of half(clamp(h)), with
// d_q(y) = base[ROWB * y + q] (the base population's deviation from w_q), and loads h = (float)half + d_q(y); the
// subtraction and the addition are FP32 operations, each rounded once.  The memory layout is the plain format's.
template <int FMT> __host__ __device__ constexpr bool row_based() { return FMT == FMT_F16R || FMT == FMT_BF16R; }
constexpr int ROWB = 16; // floats per row of a base table (9 used)

typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)nbytes, 0x00020000);
}

// Register <-> storage conversions of one population (the storage word holds f for FMT_F32, h otherwise).
// The 16-bit formats also convert two populations at once (st2: one packed conversion instruction, the
// same round-to-nearest-even as the single form; the word's low half is population q0, the high half q1).
template <int FMT> struct Conv;
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

template <> struct Conv<FMT_F32> {
    using T = float;
    static constexpr unsigned BYTES = 4;
    __device__ static float ld(T v, int q) { return v - wq(q); }
    __device__ static T st(float h, int q) { return clampf(wq(q) + h, 0.0f, 10.0f); }
    __device__ static T st(float h, int q, const CollideK &) { return st(h, q); }
    __device__ static float pop(T v, int) { return v; } // the population f itself
    __device__ static T from_pop(float f, int) { return f; }
};

template <> struct Conv<FMT_F16> {
    using T = uint16_t;
    static constexpr unsigned BYTES = 2;
    __device__ static float ld(T v, int) { return (float)__builtin_bit_cast(_Float16, v); }
    __device__ static T st(float h, int q)
    {
        return __builtin_bit_cast(uint16_t, (_Float16)clampf(h, -wq(q), 10.0f - wq(q)));
    }
    __device__ static T st(float h, int q, const CollideK &K)
    {
        return __builtin_bit_cast(uint16_t, (_Float16)clampf(h, -wq(q), K.hi[wclass(q)]));
    }
    __device__ static unsigned st2(float h0, int q0, float h1, int q1, const CollideK &K)
    {
        const f32x2_t v = {clampf(h0, -wq(q0), K.hi[wclass(q0)]), clampf(h1, -wq(q1), K.hi[wclass(q1)])};
        return __builtin_bit_cast(unsigned, __builtin_convertvector(v, f16x2_t));
    }
    __device__ static float pop(T v, int q) { return wq(q) + ld(v, q); }
    __device__ static T from_pop(float f, int q) { return __builtin_bit_cast(uint16_t, (_Float16)(f - wq(q))); }
};

template <> struct Conv<FMT_BF16> {
    using T = uint16_t;
    static constexpr unsigned BYTES = 2;
    __device__ static float ld(T v, int) { return __builtin_bit_cast(float, (unsigned)v << 16); }
    __device__ static T st(float h, int q)
    {
        return __builtin_bit_cast(uint16_t, (__bf16)clampf(h, -wq(q), 10.0f - wq(q)));
    }
    __device__ static T st(float h, int q, const CollideK &K)
    {
        return __builtin_bit_cast(uint16_t, (__bf16)clampf(h, -wq(q), K.hi[wclass(q)]));
    }
    __device__ static unsigned st2(float h0, int q0, float h1, int q1, const CollideK &K)
    {
        const f32x2_t v = {clampf(h0, -wq(q0), K.hi[wclass(q0)]), clampf(h1, -wq(q1), K.hi[wclass(q1)])};
        return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_t));
    }
    __device__ static float pop(T v, int q) { return wq(q) + ld(v, q); }
    __device__ static T from_pop(float f, int q) { return __builtin_bit_cast(uint16_t, (__bf16)(f - wq(q))); }
};

// The row-based formats: the plain format's conversions plus their base-relative forms.  ldb adds the base value d to
// the stored deviation (FP32, one rounding); with `one` = 1.0 held in a VGPR the FP16 form is one v_fma_mix_f32 (the
// half converted exactly, then one rounding: the bits of v_cvt_f32_f16 + v_add_f32), which the compiler would
// otherwise rewrite as a conversion and an add.  stb clamps h like st, subtracts d (FP32, one rounding) and rounds
// the difference to 16 bits.  (Rounding the exact difference once instead -- one v_fma_mixlo_f16, which is what the
// compiler makes of a 16-bit rounding of an FP32 fma -- gives other bits in rare ties and measured slower.)
template <int PLAIN> struct ConvRow : Conv<PLAIN> {
    using P = Conv<PLAIN>;
    using T = typename P::T;
    __device__ static float ldb(T v, float d) { return P::ld(v, 0) + d; }
    __device__ static float ldb(T v, float d, float one)
    {
        if constexpr (PLAIN == FMT_F16)
            return fma_(P::ld(v, 0), one, d);
        else
            return P::ld(v, 0) + d;
    }
    __device__ static float clamp_h(float h, int q, const CollideK &K) { return clampf(h, -wq(q), K.hi[wclass(q)]); }
    __device__ static T rnd(float v)
    {
        if constexpr (PLAIN == FMT_F16)
            return __builtin_bit_cast(uint16_t, (_Float16)v);
        else
            return __builtin_bit_cast(uint16_t, (__bf16)v);
    }
    __device__ static T stb(float h, int q, float d, const CollideK &K) { return rnd(clamp_h(h, q, K) - d); }
    __device__ static T stb(float h, int q, float d) { return stb(h, q, d, CollideK{}); }
    __device__ static unsigned st2b(float h0, int q0, float d0, float h1, int q1, float d1, const CollideK &K)
    {
        const f32x2_t v = {clamp_h(h0, q0, K) - d0, clamp_h(h1, q1, K) - d1};
        if constexpr (PLAIN == FMT_F16)
            return __builtin_bit_cast(unsigned, __builtin_convertvector(v, f16x2_t));
        else
            return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_t));
    }
    // populations f = w + (stored + d) and back (no clamp), for conversions and macro fields
    __device__ static float popb(T v, int q, float d) { return wq(q) + ldb(v, d); }
    __device__ static T from_popb(float f, int q, float d) { return rnd((f - wq(q)) - d); }
};
template <> struct Conv<FMT_F16R> : ConvRow<FMT_F16> {};
template <> struct Conv<FMT_BF16R> : ConvRow<FMT_BF16> {};

// Buffer-resource loads and stores of one population in storage format FMT.
template <int FMT> struct Store {
    using C = Conv<FMT>;
    static constexpr unsigned BYTES = C::BYTES;
    __device__ static float ld(rsrc_t r, unsigned voff, unsigned soff, int q)
    {
        if constexpr (BYTES == 4)
            return C::ld(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0)), q);
        else
            return C::ld((uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0), q);
    }
    __device__ static void st(rsrc_t r, unsigned voff, unsigned soff, int q, float h)
    {
        if constexpr (BYTES == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, C::st(h, q)), r, voff, soff, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b16(C::st(h, q), r, voff, soff, 0);
    }
    // row-based formats: load adding, store subtracting the base value d of the slot
    __device__ static float ldb(rsrc_t r, unsigned voff, unsigned soff, float d, float one)
    {
        return C::ldb((uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0), d, one);
    }
    __device__ static void stb(rsrc_t r, unsigned voff, unsigned soff, int q, float h, float d)
    {
        __builtin_amdgcn_raw_buffer_store_b16(C::stb(h, q, d), r, voff, soff, 0);
    }
};

// Base values of one row (the 9 used entries of a base table row) in scalar registers: y must be the same in every
// active lane of the wave (it is read from the first one); the table is read through the constant address space,
// so the loads are scalar (two per row).
typedef const __attribute__((address_space(4))) float *const_fptr;
__device__ __forceinline__ void base_row_s(const float *base, int y, float d[9])
{
    const int yu = __builtin_amdgcn_readfirstlane(y);
    const const_fptr t = (const_fptr)base + ROWB * yu;
#pragma unroll
    for (int q = 0; q < 9; ++q) d[q] = t[q];
}

// The base window of the row-based formats: whether cell (x, y) lies in it, and its base values there
__device__ __forceinline__ bool in_window(const KPar &P, int x, int y)
{
    return (unsigned)(x - P.wx0) < (unsigned)P.ww && (unsigned)(y - P.wy0) < (unsigned)P.wh;
}
// packed velocity of window cell (x, y), or of the nearest window cell when (x, y) lies outside it (a load from a
// valid address either way, so that a wave's loads issue together; needs a non-empty window)
__device__ __forceinline__ uint32_t window_vel(const KPar &P, int x, int y)
{
    const int ix = min(max(x - P.wx0, 0), P.ww - 1), iy = min(max(y - P.wy0, 0), P.wh - 1);
    return P.wvel[(unsigned)(iy * P.ww + ix)];
}
// (ux, uy) of a packed window velocity
__device__ __forceinline__ void unpack_vel(uint32_t v, float &ux, float &uy)
{
    ux = (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
    uy = (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
}
// The window's base value of direction q: the equilibrium deviation at rho = 1 of velocity (ux, uy),
// w_q ((3 cu + 4.5 cu^2) - 1.5 |u|^2) with cu = c_q.u (one addition for the diagonals, exact for the axes), in this
// order of FP32 operations everywhere (opposite directions give -cu exactly).  usq15 = 1.5 (ux ux + uy uy).
__device__ __forceinline__ float vel_usq15(float ux, float uy) { return 1.5f * (ux * ux + uy * uy); }
__device__ __forceinline__ float vel_base(int q, float ux, float uy, float usq15)
{
    const float cu = q == 0 ? 0.0f
                            : (cx(q) == 0 ? (cy(q) > 0 ? uy : -uy)
                                          : (cy(q) == 0 ? (cx(q) > 0 ? ux : -ux)
                                                        : (cx(q) > 0 ? ux : -ux) + (cy(q) > 0 ? uy : -uy)));
    return wq(q) * ((3.0f * cu + 4.5f * (cu * cu)) - usq15);
}
// base value q of cell (x, y) where the cell lies in the window, else `other` (branch-free; global table)
__device__ __forceinline__ float window_or(const KPar &P, int x, int y, int q, float other)
{
    float ux, uy;
    unpack_vel(window_vel(P, x, y), ux, uy);
    const float v = vel_base(q, ux, uy, vel_usq15(ux, uy));
    return in_window(P, x, y) ? v : other;
}
// whether the rectangle [x0, x1] x [y0, y1] meets the window
__device__ __forceinline__ bool meets_window(const KPar &P, int x0, int x1, int y0, int y1)
{
    return P.ww > 0 && x1 >= P.wx0 && x0 < P.wx0 + P.ww && y1 >= P.wy0 && y0 < P.wy0 + P.wh;
}

// 1.0 in a VGPR the compiler cannot see through (the multiplier of the FP16 ldb form)
__device__ __forceinline__ float opaque_one()
{
    float one = 1.0f;
    asm volatile("" : "+v"(one));
    return one;
}

// FP32 populations in registers (pull kernel and the fused kernel's LDS): f -> h -> collide -> f'
template <int MODEL>
__device__ __forceinline__ void collide(const KPar &P, float f[9], bool inlet, float uin, float mac[3])
{
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = f[q] - wq(q);
    collide_dev<MODEL>(P, f, inlet, uin, mac);
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = clampf(wq(q) + f[q], 0.0f, 10.0f);
}

} // namespace lbm
