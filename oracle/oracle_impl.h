/*
 * CPU restatement of the D2Q9 collide-stream step -- TEST INFRASTRUCTURE ONLY.
 *
 * This header is included twice by lbm_oracle.c: once with REAL = double (the FP64 oracle) and once with
 * REAL = float (the FP32 mirror, whose operation order and fused multiply-adds are the ones the gfx950
 * kernels use, so that the two agree bit for bit).  Nothing in the product links or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, as the checker.
 *
 * What it restates (file:line in the reference tree):
 *   - D2Q9 velocity set, opposite table, weights           lbm_core4.cu:1-6, lbm_core_smemaa.cu:1-7
 *   - moments, clamps, MRT forward/equilibria/relax/inverse lbm_core4.cu:124-199
 *   - Smagorinsky tau_eff                                   lbm_core4.cu:171-177
 *   - population clamp [0,10] and outlet ghost copy         lbm_core4.cu:200-210
 *   - Zou-He velocity inlet + inlet MRT (no LES, no clamps) lbm_core4.cu:29-118
 *   - pull streaming with half-way bounce-back              lbm_core4.cu:129-138
 *   - BGK collision of the temporally blocked family        lbm_core_smemaa.cu:194-220
 *   - AA even (local, canonical->reversed) and odd (gather/scatter, reversed->canonical) steps,
 *     bounce-back on solid source/destination, left-open inlet drop
 *                                                           lbm_core_smemaa.cu:92-223, 293-445;
 *                                                           lbm_core_smemaa_endfused.cu:330-369
 *   - macro extraction with Zou-He at x=0                   lbm_core_smemaa_endfused.cu:377-422
 *
 * Deliberate definitions where the reference is inconsistent (each one is pinned by a test):
 *   - corners (W-1,0) and (W-1,H-1) are walls (lbm_core4.cu:134 honours the mask; the AA kernels force the
 *     whole W-1 column open, lbm_core_smemaa.cu:51-55);
 *   - the zero-gradient outlet is the one of lbm_core4.cu:114-116 (ghost W-1 = post-collision copy of W-2).
 *     In AA form this needs the W-2 cell to stand in for the ghost's scatter of q=3,6,7 on odd steps, which
 *     the reference's AA kernels omit (their W-2 canonical slots 3,6,7 go stale, lbm_core_smemaa.cu:296-315);
 *   - the inlet speed is a per-row array (the reference evaluates u_in*ramp(y) inline, lbm_core4.cu:46-51).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUFFIX)

/* ---------------------------------------------------------------------------------------------------- */
/* Derived constants                                                                                      */
/* ---------------------------------------------------------------------------------------------------- */

typedef struct {
    int W, H, model, bc, clamp, les, sums;
    REAL tau_inv, tau0, tau0sq, c18cs2, s_e, s_eps, s_q;
} FN(kpar_t);

static void FN(make_kpar)(const oracle_params *p, FN(kpar_t) * k)
{
    k->W = p->W;
    k->H = p->H;
    k->model = p->model;
    k->bc = p->bc;
    k->clamp = p->clamp;
    k->sums = p->sums;
    k->les = (p->model == ORACLE_MRT && p->cs != 0.0f) ? 1 : 0;
    k->tau_inv = (REAL)p->tau_inv;
    /* tau_0 = 1/tau_inv and 18*Cs*Cs as in lbm_core4.cu:175-176, evaluated once in the working precision */
    k->tau0 = (REAL)1 / (REAL)p->tau_inv;
    k->tau0sq = k->tau0 * k->tau0;
    k->c18cs2 = ((REAL)18 * (REAL)p->cs) * (REAL)p->cs;
    k->s_e = (REAL)p->s_e;
    k->s_eps = (REAL)p->s_eps;
    k->s_q = (REAL)p->s_q;
}

static inline REAL FN(clampr)(REAL v, REAL lo, REAL hi)
{
    /* fminf(hi, fmaxf(lo, v)) of lbm_core4.cu:142-144,203 */
    REAL a = v > lo ? v : lo;
    return a < hi ? a : hi;
}

/* ---------------------------------------------------------------------------------------------------- */
/* Collision K: 9 pre-collision populations -> 9 post-collision populations (clamped to [0,10]).          */
/* inlet != 0 applies Zou-He with speed uin first and then the inlet collision of lbm_core4.cu:65-93.       */
/* mac (may be NULL) receives (rho, ux, uy) as the reference writes them (lbm_core4.cu:59-61,145-147).     */
/* ---------------------------------------------------------------------------------------------------- */

/* Deviation form: h_q = f_q - w_q (rest state rho = 1, u = 0, whose moments 1, -2, 1, 0.. are exact). */
static inline REAL FN(wq)(int q) { return q == 0 ? C_W0 : (q < 5 ? C_W1 : C_W5); }

/* Zou-He (lbm_core4.cu:54-57) on deviations; returns rho - 1 and sets *rho. */
static inline REAL FN(zou_he)(REAL *h, REAL uin, REAL *rho)
{
    REAL Hs = FMA((REAL)2, (h[3] + h[6]) + h[7], (h[0] + h[2]) + h[4]);
    REAL drho = (Hs + uin) / ((REAL)1 - uin);
    *rho = (REAL)1 + drho;
    REAL ru = *rho * uin;
    REAL hh = (REAL)0.5 * (h[2] - h[4]);
    h[1] = FMA(C_2_3, ru, h[3]);
    h[5] = FMA(C_1_6, ru, h[7] - hh);
    h[8] = FMA(C_1_6, ru, h[6] + hh);
    return drho;
}

/* Neumaier's compensated summation (Neumaier 1974, the improved Kahan-Babuska scheme): s + t rounded, the
   rounding error of that addition accumulated in c; the sum is s + c at the end. */
static inline void FN(neu_add)(REAL *s, REAL *c, REAL t)
{
    REAL u = *s + t;
    *c += FABS(*s) >= FABS(t) ? (*s - u) + t : (t - u) + *s;
    *s = u;
}

/* rho - 1, jx, jy as compensated sums over the deviations, terms in index order */
static void FN(moments_neu)(const REAL *h, REAL *drho, REAL *jx, REAL *jy)
{
    REAL s = h[0], c = 0;
    for (int q = 1; q < 9; ++q) FN(neu_add)(&s, &c, h[q]);
    *drho = s + c;
    s = h[1], c = 0;
    FN(neu_add)(&s, &c, -h[3]);
    FN(neu_add)(&s, &c, h[5]);
    FN(neu_add)(&s, &c, -h[6]);
    FN(neu_add)(&s, &c, -h[7]);
    FN(neu_add)(&s, &c, h[8]);
    *jx = s + c;
    s = h[2], c = 0;
    FN(neu_add)(&s, &c, -h[4]);
    FN(neu_add)(&s, &c, h[5]);
    FN(neu_add)(&s, &c, h[6]);
    FN(neu_add)(&s, &c, -h[7]);
    FN(neu_add)(&s, &c, -h[8]);
    *jy = s + c;
}

/* w_ext >= 0: the stress moments relax at w_ext (an external closure's rate) instead of the model's own */
static void FN(collide_mrt)(const FN(kpar_t) * P, REAL h[9], int inlet, REAL uin, REAL *mac, REAL w_ext)
{
    REAL rho = 0, ux = 0, uy = 0, drc = 0, rc = 0;
    int les = 0;
    if (inlet) {
        drc = FN(zou_he)(h, uin, &rho);
        rc = rho;
        ux = uin;
        uy = (REAL)0;
    }
    REAL s13 = h[1] + h[3], s24 = h[2] + h[4], s57 = h[5] + h[7], s68 = h[6] + h[8];
    REAL h14 = s13 + s24, h58 = s57 + s68;
    REAL a13 = h[1] - h[3], a24 = h[2] - h[4], a57 = h[5] - h[7], a86 = h[8] - h[6];
    REAL sx = a57 + a86; /* f5 - f6 - f7 + f8 */
    REAL sy = a57 - a86; /* f5 + f6 - f7 - f8 */
    if (!inlet) {
        /* moments and clamps, lbm_core4.cu:130-144 (rho = 1 + drho) */
        REAL drho = (h[0] + h14) + h58;
        REAL jx = a13 + sx;
        REAL jy = a24 + sy;
        if (P->sums) FN(moments_neu)(h, &drho, &jx, &jy);
        rho = (REAL)1 + drho;
        REAL inv_rho = (REAL)1 / rho;
        ux = jx * inv_rho;
        uy = jy * inv_rho;
        drc = drho;
        if (P->clamp) {
            ux = FN(clampr)(ux, (REAL)-0.4, (REAL)0.4);
            uy = FN(clampr)(uy, (REAL)-0.4, (REAL)0.4);
            drc = FN(clampr)(drho, (REAL)-1, (REAL)0.5); /* rho in [0, 1.5] */
        }
        rc = (REAL)1 + drc;
        les = P->les;
    }
    if (mac) {
        mac[0] = rc;
        mac[1] = ux;
        mac[2] = uy;
    }
    /* MRT forward, Lallemand-Luo order (rho, e, eps, jx, qx, jy, qy, pxx, pxy), lbm_core4.cu:152-165;
       e = -2 + m1d, eps = 1 + m2d */
    REAL m1d = FMA((REAL)2, h58, FMA((REAL)-4, h[0], -h14));
    REAL m2d = FMA((REAL)-2, h14, FMA((REAL)4, h[0], h58));
    REAL m3 = rc * ux;
    REAL m4 = FMA((REAL)-2, a13, sx);
    REAL m5 = rc * uy;
    REAL m6 = FMA((REAL)-2, a24, sy);
    REAL m7 = s13 - s24;
    REAL m8 = s57 - s68;
    /* non-equilibrium parts, lbm_core4.cu:166-170 */
    REAL ux2 = ux * ux, uy2 = uy * uy;
    REAL usq = ux2 + uy2;
    REAL u3 = (REAL)3 * usq;
    REAL n1 = FMA(-u3, rc, FMA((REAL)2, drc, m1d)); /* e - (3u^2 - 2) rho */
    REAL n2 = FMA(u3, rc, m2d - drc);              /* eps - (1 - 3u^2) rho */
    REAL m7eq = rc * (ux2 - uy2);
    REAL m8eq = m3 * uy;
    REAL n7 = m7 - m7eq;
    REAL n8 = m8 - m8eq;
    REAL w = P->tau_inv;
    if (w_ext >= (REAL)0) {
        w = w_ext;
        les = 0;
    }
    if (les) {
        /* Smagorinsky, lbm_core4.cu:172-177 */
        REAL Q = SQRT(FMA(n7, n7, FMA(n8, n8, (REAL)1e-8)));
        REAL inv_rc = (REAL)1 / rc;
        REAL d = FMA(P->c18cs2, Q * inv_rc, P->tau0sq);
        REAL tau_eff = (REAL)0.5 * (P->tau0 + SQRT(d));
        w = (REAL)1 / tau_eff;
    }
    /* relaxation, lbm_core4.cu:180-185 (m4, m6 relax towards -m3, -m5) */
    m1d = FMA(-P->s_e, n1, m1d);
    m2d = FMA(-P->s_eps, n2, m2d);
    m4 = FMA(-P->s_q, m4 + m3, m4);
    m6 = FMA(-P->s_q, m6 + m5, m6);
    m7 = FMA(-w, n7, m7);
    m8 = FMA(-w, n8, m8);
    /* inverse, lbm_core4.cu:187-199, of the deviation moments (rho - 1, m1d, m2d, m3, ..., m8), with the
       common 1/36 folded into the coefficients: 6/36 = 1/6, 3/36 = 1/12, 9/36 = 1/4 */
    REAL t0 = (drc - m1d) + m2d;
    REAL t1 = FMA((REAL)-2, m2d, FMA((REAL)4, drc, -m1d));
    REAL t2 = FMA((REAL)4, drc, FMA((REAL)2, m1d, m2d));
    REAL q7 = m7 * (REAL)0.25;
    REAL q8 = m8 * (REAL)0.25;
    REAL t1p = FMA(t1, C_INV36, q7), t1m = FMA(t1, C_INV36, -q7);
    REAL d34 = m3 - m4, d56 = m5 - m6;
    REAL X = FMA(C_1_6, m3, m4 * C_1_12);
    REAL Y = FMA(C_1_6, m5, m6 * C_1_12);
    REAL u = FMA(t2, C_INV36, q8), v = FMA(t2, C_INV36, -q8);
    REAL XpY = X + Y, XmY = X - Y;
    h[0] = t0 * C_INV9;
    h[1] = FMA(d34, C_1_6, t1p);
    h[2] = FMA(d56, C_1_6, t1m);
    h[3] = FMA(d34, -C_1_6, t1p);
    h[4] = FMA(d56, -C_1_6, t1m);
    h[5] = u + XpY;
    h[6] = v - XmY;
    h[7] = u - XpY;
    h[8] = v + XmY;
}

/* BGK of lbm_core_smemaa.cu:194-212 on deviations: heq_q = w_q (drho + rho (3cu + 4.5cu^2 - 1.5u^2)) */
static void FN(collide_bgk)(const FN(kpar_t) * P, REAL h[9], int inlet, REAL uin, REAL *mac)
{
    REAL rho, drho, ux, uy;
    if (inlet) {
        drho = FN(zou_he)(h, uin, &rho);
        ux = uin;
        uy = (REAL)0;
    } else {
        REAL s13 = h[1] + h[3], s24 = h[2] + h[4], s57 = h[5] + h[7], s68 = h[6] + h[8];
        REAL a13 = h[1] - h[3], a24 = h[2] - h[4], a57 = h[5] - h[7], a86 = h[8] - h[6];
        drho = (h[0] + (s13 + s24)) + (s57 + s68);
        rho = (REAL)1 + drho;
        REAL inv_rho = (REAL)1 / rho;
        ux = (a13 + (a57 + a86)) * inv_rho;
        uy = (a24 + (a57 - a86)) * inv_rho;
    }
    if (mac) {
        mac[0] = rho;
        mac[1] = ux;
        mac[2] = uy;
    }
    REAL usq15 = (REAL)1.5 * FMA(uy, uy, ux * ux);
    REAL cu[9];
    cu[0] = (REAL)0;
    cu[1] = ux;
    cu[2] = uy;
    cu[3] = -ux;
    cu[4] = -uy;
    cu[5] = ux + uy;
    cu[6] = uy - ux;
    cu[7] = -(ux + uy);
    cu[8] = ux - uy;
    for (int q = 0; q < 9; ++q) {
        REAL poly = q == 0 ? -usq15 : FMA(cu[q], FMA((REAL)4.5, cu[q], (REAL)3), -usq15);
        REAL heq = FN(wq)(q) * FMA(rho, poly, drho);
        h[q] = FMA(P->tau_inv, heq - h[q], h[q]);
    }
}

/* Collision of FP32/FP64-stored populations: h = f - w, collide, f' = clamp(w + h', 0, 10). */
static inline void FN(collide)(const FN(kpar_t) * P, REAL f[9], int inlet, REAL uin, REAL *mac, REAL w_ext)
{
    for (int q = 0; q < 9; ++q) f[q] = f[q] - FN(wq)(q);
    if (P->model == ORACLE_BGK)
        FN(collide_bgk)(P, f, inlet, uin, mac);
    else
        FN(collide_mrt)(P, f, inlet, uin, mac, w_ext);
    for (int q = 0; q < 9; ++q) f[q] = FN(clampr)(FN(wq)(q) + f[q], (REAL)0, (REAL)10);
}

/* ---------------------------------------------------------------------------------------------------- */
/* Streaming                                                                                              */
/* ---------------------------------------------------------------------------------------------------- */

/* neighbour of (x,y) displaced by (dx,dy); returns -1 when it leaves a channel domain */
static inline long FN(nbr)(const FN(kpar_t) * P, int x, int y, int dx, int dy)
{
    int nx = x + dx, ny = y + dy;
    if (P->bc == ORACLE_PERIODIC) {
        if (nx < 0) nx += P->W;
        if (nx >= P->W) nx -= P->W;
        if (ny < 0) ny += P->H;
        if (ny >= P->H) ny -= P->H;
    } else if (nx < 0 || nx >= P->W || ny < 0 || ny >= P->H) {
        return -1;
    }
    return (long)ny * P->W + nx;
}

/* 0 = not updated (solid, wall row, outlet ghost), 1 = fluid, 2 = inlet   (lbm_core4.cu:20-29,121) */
static inline int FN(role)(const FN(kpar_t) * P, const uint8_t *mask, int x, int y)
{
    if (P->bc == ORACLE_PERIODIC) return mask[(long)y * P->W + x] ? 0 : 1;
    if (y <= 0 || y >= P->H - 1 || x >= P->W - 1) return 0;
    if (x == 0) return 2;
    return mask[(long)y * P->W + x] ? 0 : 1;
}

/* The incoming populations of cell (x, y) of an AA state of the given parity (see oracle_aa_step). */
static void FN(gather)(const FN(kpar_t) * P, const uint8_t *mask, const REAL *f, int parity, int x, int y,
                       int inlet, REAL g[9])
{
    const long plane = (long)P->W * P->H;
    const long idx = (long)y * P->W + x;
    if (parity == 0) {
        for (int q = 0; q < 9; ++q) g[q] = f[q * plane + idx];
        return;
    }
    for (int q = 0; q < 9; ++q) {
        if (inlet && CX[q] == 1) {
            g[q] = (REAL)0; /* from x = -1: rebuilt by Zou-He */
            continue;
        }
        long s = FN(nbr)(P, x, y, -CX[q], -CY[q]);
        g[q] = mask[s] ? f[q * plane + idx] : f[OPP[q] * plane + s];
    }
}

static void FN(aa_step_impl)(const oracle_params *pp, const uint8_t *mask, REAL *f, const REAL *inlet_u,
                             int parity, REAL *oux, REAL *ouy, REAL *orho, const REAL *w_cell);

/* One AA step in place.  parity 0: canonical -> reversed (local).  parity 1: reversed -> canonical. */
void FN(oracle_aa_step)(const oracle_params *pp, const uint8_t *mask, REAL *f, const REAL *inlet_u, int parity,
                        REAL *oux, REAL *ouy, REAL *orho)
{
    FN(aa_step_impl)(pp, mask, f, inlet_u, parity, oux, ouy, orho, NULL);
}

/* oracle_aa_step whose MRT stress moments relax at w_cell[cell] where that is >= 0 (the learned closure's
   rate); cells with w_cell < 0 use the model's own rule. */
void FN(oracle_aa_step_rate)(const oracle_params *pp, const uint8_t *mask, REAL *f, const REAL *inlet_u,
                             int parity, const REAL *w_cell)
{
    FN(aa_step_impl)(pp, mask, f, inlet_u, parity, NULL, NULL, NULL, w_cell);
}

/* m7neq and m8neq of the pre-collision state of every updated cell (collide_mrt's n7, n8) into
   neq[0][cell], neq[1][cell]; other cells get 0. */
void FN(oracle_neq)(const oracle_params *pp, const uint8_t *mask, const REAL *f, const REAL *inlet_u, int parity,
                    REAL *neq)
{
    FN(kpar_t) P;
    FN(make_kpar)(pp, &P);
    const long plane = (long)P.W * P.H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < P.H; ++y) {
        for (int x = 0; x < P.W; ++x) {
            const long idx = (long)y * P.W + x;
            neq[idx] = neq[plane + idx] = (REAL)0;
            int r = FN(role)(&P, mask, x, y);
            if (!r) continue;
            REAL g[9];
            FN(gather)(&P, mask, f, parity, x, y, r == 2, g);
            REAL s13, s24, s57, s68, h[9];
            for (int q = 0; q < 9; ++q) h[q] = g[q] - FN(wq)(q);
            s13 = h[1] + h[3], s24 = h[2] + h[4], s57 = h[5] + h[7], s68 = h[6] + h[8];
            REAL h14 = s13 + s24, h58 = s57 + s68;
            REAL a13 = h[1] - h[3], a24 = h[2] - h[4], a57 = h[5] - h[7], a86 = h[8] - h[6];
            REAL sx = a57 + a86, sy = a57 - a86;
            REAL drho = (h[0] + h14) + h58;
            REAL rho = (REAL)1 + drho;
            REAL inv_rho = (REAL)1 / rho;
            REAL ux = (a13 + sx) * inv_rho, uy = (a24 + sy) * inv_rho, drc = drho;
            if (P.clamp) {
                ux = FN(clampr)(ux, (REAL)-0.4, (REAL)0.4);
                uy = FN(clampr)(uy, (REAL)-0.4, (REAL)0.4);
                drc = FN(clampr)(drho, (REAL)-1, (REAL)0.5);
            }
            REAL rc = (REAL)1 + drc;
            REAL m3 = rc * ux;
            neq[idx] = (s13 - s24) - rc * (ux * ux - uy * uy);
            neq[plane + idx] = (s57 - s68) - m3 * uy;
        }
    }
}

static void FN(aa_step_impl)(const oracle_params *pp, const uint8_t *mask, REAL *f, const REAL *inlet_u,
                             int parity, REAL *oux, REAL *ouy, REAL *orho, const REAL *w_cell)
{
    FN(kpar_t) P;
    FN(make_kpar)(pp, &P);
    const long plane = (long)P.W * P.H;
    const int channel = P.bc == ORACLE_CHANNEL;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < P.H; ++y) {
        for (int x = 0; x < P.W; ++x) {
            int r = FN(role)(&P, mask, x, y);
            if (!r) continue;
            const int inlet = r == 2;
            const long idx = (long)y * P.W + x;
            REAL g[9], mac[3];
            FN(gather)(&P, mask, f, parity, x, y, inlet, g);
            FN(collide)(&P, g, inlet, inlet ? inlet_u[y] : (REAL)0, mac, w_cell ? w_cell[idx] : (REAL)-1);
            if (parity == 0) {
                for (int q = 0; q < 9; ++q) f[OPP[q] * plane + idx] = g[q];
                if (channel && x == P.W - 2)
                    for (int q = 0; q < 9; ++q) f[OPP[q] * plane + idx + 1] = g[q];
            } else {
                for (int q = 0; q < 9; ++q) {
                    if (inlet && CX[q] == -1) continue; /* leaves through the open inlet */
                    long d = FN(nbr)(&P, x, y, CX[q], CY[q]);
                    if (mask[d])
                        f[OPP[q] * plane + idx] = g[q];
                    else
                        f[q * plane + d] = g[q];
                }
                if (channel && x == P.W - 2) {
                    /* the ghost (W-1,y) holds a copy of this cell; stream its q=3,6,7 into column W-2 */
                    for (int q = 0; q < 9; ++q) {
                        if (CX[q] != -1) continue;
                        long d = (long)(y + CY[q]) * P.W + x;
                        if (!mask[d]) f[q * plane + d] = g[q];
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

/* Reference-compatible pull step (lbm_core4.cu:120-211): f_out holds post-collision populations. */
void FN(oracle_pull_step)(const oracle_params *pp, const uint8_t *mask, const REAL *f_now, REAL *f_out,
                          const REAL *inlet_u, REAL *oux, REAL *ouy, REAL *orho)
{
    FN(kpar_t) P;
    FN(make_kpar)(pp, &P);
    const long plane = (long)P.W * P.H;
    const int channel = P.bc == ORACLE_CHANNEL;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < P.H; ++y) {
        for (int x = 0; x < P.W; ++x) {
            int r = FN(role)(&P, mask, x, y);
            if (!r) continue;
            const int inlet = r == 2;
            const long idx = (long)y * P.W + x;
            REAL g[9], mac[3];
            for (int q = 0; q < 9; ++q) {
                if (inlet && CX[q] == 1) {
                    g[q] = (REAL)0;
                    continue;
                }
                long s = FN(nbr)(&P, x, y, -CX[q], -CY[q]);
                g[q] = mask[s] ? f_now[OPP[q] * plane + idx] : f_now[q * plane + s];
            }
            FN(collide)(&P, g, inlet, inlet ? inlet_u[y] : (REAL)0, mac, (REAL)-1);
            for (int q = 0; q < 9; ++q) {
                f_out[q * plane + idx] = g[q];
                if (channel && x == P.W - 2) f_out[q * plane + idx + 1] = g[q];
            }
            if (orho) {
                orho[idx] = mac[0];
                oux[idx] = mac[1];
                ouy[idx] = mac[2];
            }
        }
    }
}

/* rho, u of the stored state (lbm_core_smemaa_endfused.cu:377-422): canonical states get Zou-He at x=0,
 * reversed states are read through opp().  Cells that are not updated are left untouched. */
void FN(oracle_macro)(const oracle_params *pp, const uint8_t *mask, const REAL *f, const REAL *inlet_u, int parity,
                      REAL *oux, REAL *ouy, REAL *orho)
{
    FN(kpar_t) P;
    FN(make_kpar)(pp, &P);
    const long plane = (long)P.W * P.H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < P.H; ++y) {
        for (int x = 0; x < P.W; ++x) {
            int r = FN(role)(&P, mask, x, y);
            if (!r) continue;
            const long idx = (long)y * P.W + x;
            REAL g[9];
            for (int q = 0; q < 9; ++q) g[q] = f[(parity ? OPP[q] : q) * plane + idx];
            REAL rho, ux, uy;
            if (r == 2 && parity == 0) {
                REAL u = inlet_u[y];
                rho = FMA((REAL)2, (g[3] + g[6]) + g[7], (g[0] + g[2]) + g[4]) / ((REAL)1 - u);
                ux = u;
                uy = (REAL)0;
            } else {
                REAL s13 = g[1] + g[3], s24 = g[2] + g[4], s57 = g[5] + g[7], s68 = g[6] + g[8];
                REAL a13 = g[1] - g[3], a24 = g[2] - g[4], a57 = g[5] - g[7], a86 = g[8] - g[6];
                rho = (g[0] + (s13 + s24)) + (s57 + s68);
                ux = (a13 + (a57 + a86)) / rho;
                uy = (a24 + (a57 - a86)) / rho;
            }
            orho[idx] = rho;
            oux[idx] = ux;
            ouy[idx] = uy;
        }
    }
}

/* Exposed collision on one cell, for unit tests. */
void FN(oracle_collide)(const oracle_params *pp, REAL *f9, int inlet, REAL uin, REAL *mac3)
{
    FN(kpar_t) P;
    FN(make_kpar)(pp, &P);
    FN(collide)(&P, f9, inlet, uin, mac3, (REAL)-1);
}

#undef FN
#undef CAT
#undef CAT_
