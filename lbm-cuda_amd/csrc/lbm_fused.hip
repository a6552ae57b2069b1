// C-ABI of the temporally blocked kernel: lbm_step_fused / lbm_step_fused_cfg (FP32 populations) and
// lbm_step_fused_h16 (16-bit deviations).  The kernels live in lbm_fused_impl.h and are instantiated per
// storage format in lbm_fused_{f32,f16,bf16}.hip.
#include "lbm_common.h"

namespace lbm {

template <int FMT>
int launch_fused_fmt(const KPar &P, int model, int bc, int k, int strip, int parity, hipStream_t st,
                     const uint8_t *mask, const void *fin, void *fout, const float *inlet_u, int CH, float *ux,
                     float *uy, float *rho);

// defined in lbm_pair_impl.h (instantiated in lbm_pair.hip and lbm_pair_rb.hip)
template <int FMT>
int launch_pair_fmt(const KPar &P, int model, int bc, int CH, hipStream_t st, const uint8_t *mask, const void *fin,
                    void *fout, const float *inlet_u, float *ux, float *uy, float *rho);
// defined in lbm_quad.hip (FP32 populations)
template <int FMT>
int launch_quad_fmt(const KPar &P, int model, int bc, int CH, hipStream_t st, const uint8_t *mask, const void *fin,
                    void *fout, const float *inlet_u, float *ux, float *uy, float *rho);

// strip = 0 selects a tile kernel from an even state -- the two-step kernel for k = 2 (chunk = 8, 12, 16 or 24
// tile rows), the four-step kernel for k = 4 (chunk = 8 or 16 rows of 128-wide tiles, 12, 14, 20, 24 or 32 rows
// of 64-wide ones; FP32 populations); otherwise the strip-sweep kernel
template <int FMT>
static int launch_any(const KPar &P, const lbm_params *p, int k, const lbm_fused_cfg &c, int parity, hipStream_t st,
                      const uint8_t *mask, const void *fin, void *fout, const float *inlet_u, float *ux, float *uy,
                      float *rho)
{
    if (c.strip == 0 && k == 4) {
        if constexpr (FMT == FMT_F32)
            return launch_quad_fmt<FMT>(P, p->model, p->bc, c.chunk, st, mask, fin, fout, inlet_u, ux, uy, rho);
        else
            return fail(LBM_ERR_UNSUPPORTED, "lbm_step_fused: the four-step tile kernel stores FP32 populations only");
    }
    if (c.strip == 0) return launch_pair_fmt<FMT>(P, p->model, p->bc, c.chunk, st, mask, fin, fout, inlet_u, ux, uy, rho);
    return launch_fused_fmt<FMT>(P, p->model, p->bc, k, c.strip, parity, st, mask, fin, fout, inlet_u, c.chunk, ux,
                                 uy, rho);
}

// Strip configuration offered by lbm_fused_default_cfg: 128-column strips, and the tallest chunk that still gives
// at least 512 workgroups (two per CU on 256 CUs) so the k-row warm-up at each chunk end stays small.
static void default_cfg(int W, int H, int k, lbm_fused_cfg *c)
{
    c->strip = 128;
    const int CW = 128 - 2 * (k + 1);
    const int strips = (W + CW - 1) / CW;
    int ch = 256;
    while (ch > 16 && (long long)strips * ((H + ch - 1) / ch) < 512) ch /= 2;
    c->chunk = ch < H ? ch : H;
}


// Checks the arguments and picks the launch configuration c.  Without a configuration (cfg = NULL): the tile kernel
// when one launch of it covers the k steps, otherwise `composed` (run_composed below); the strip sweep runs for an
// explicit strip configuration only.
static int fused_common(const lbm_params *p, const uint8_t *mask, const void *f_now, void *f_out, unsigned bytes,
                        const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                        float *rho, KPar &P, lbm_fused_cfg &c, bool &composed)
{
    composed = false;
    int rc = make_kpar(p, &P, bytes);
    if (rc) return rc;
    if (!mask || !f_now || !f_out) return fail(LBM_ERR_ARG, "lbm_step_fused: mask, f_now, f_out must be non-NULL");
    {
        const char *a = static_cast<const char *>(f_now), *b = static_cast<const char *>(f_out);
        const size_t n = (size_t)bytes * 9 * (size_t)P.plane;
        if (a < b + n && b < a + n) return fail(LBM_ERR_ARG, "lbm_step_fused: f_now and f_out overlap");
    }
    if (p->bc == LBM_BC_CHANNEL && !inlet_u) return fail(LBM_ERR_ARG, "lbm_step_fused: channel mode needs inlet_u");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_step_fused: parity must be 0 or 1");
    if ((ux == nullptr) != (rho == nullptr) || (uy == nullptr) != (rho == nullptr))
        return fail(LBM_ERR_ARG, "lbm_step_fused: ux, uy, rho must be all NULL or all non-NULL");
    if (k < 1 || k > FUSED_MAXK) return fail(LBM_ERR_ARG, "lbm_step_fused: k=%d outside [1, %d]", k, FUSED_MAXK);
    if (cfg) {
        c = *cfg;
    } else if (parity == 0 && (k == 2 || (k == 4 && bytes == 4))) {
        c = lbm_fused_cfg{0, 12}; // from an even state the tile kernels: 2-3x the strip sweep's rate at 4096 x 1024
    } else {
        composed = true; // every form the single-step kernel takes
        return LBM_OK;
    }
    if (P.fast && c.strip != 0)
        return fail(LBM_ERR_UNSUPPORTED, "lbm_step_fused: fast arithmetic takes the tile kernels (strip 0) or no "
                                         "configuration, not the strip sweep");
    if (P.base && bytes == 2 && !(c.strip == 0 && k == 2))
        return fail(LBM_ERR_UNSUPPORTED, "lbm_step_fused_h16: row-based storage (p->row_base) takes the two-step "
                                         "tile kernel (strip 0, k = 2) or no configuration, not the strip sweep");
    if (c.strip == 0) {
        if ((k != 2 && k != 4) || parity != 0)
            return fail(LBM_ERR_UNSUPPORTED,
                        "lbm_step_fused: the tile kernels (strip 0) need k = 2 or 4 and parity 0");
        if (k == 2 && c.chunk != 8 && c.chunk != 12 && c.chunk != 16 && c.chunk != 24)
            return fail(LBM_ERR_ARG, "lbm_step_fused: strip 0 with k = 2 needs chunk 8, 12, 16 or 24");
        if (k == 4 && c.chunk != 8 && c.chunk != 12 && c.chunk != 14 && c.chunk != 16 && c.chunk != 20 &&
            c.chunk != 24 && c.chunk != 32)
            return fail(LBM_ERR_ARG, "lbm_step_fused: strip 0 with k = 4 needs chunk 8, 12, 14, 16, 20, 24 or 32");
        if (k == 4 && bytes != 4)
            return fail(LBM_ERR_UNSUPPORTED, "lbm_step_fused: the four-step tile kernel stores FP32 populations only");
        return LBM_OK;
    }
    if (c.strip != 64 && c.strip != 128) return fail(LBM_ERR_ARG, "lbm_step_fused: strip must be 0, 64 or 128");
    if (c.chunk < 1) return fail(LBM_ERR_ARG, "lbm_step_fused: chunk must be >= 1");
    if (c.strip - 2 * (k + 1) < 16) return fail(LBM_ERR_ARG, "lbm_step_fused: strip %d too narrow for k=%d", c.strip, k);
    if (c.chunk > P.H) c.chunk = P.H;
    return LBM_OK;
}

// The first launch of lbm_step_fused_h16 for a configuration: the row-based formats run the two-step tile kernel
template <int FMT, int FMT_R>
static int launch_h16(const KPar &P, const lbm_params *p, int k, const lbm_fused_cfg &c, int parity, hipStream_t st,
                      const uint8_t *mask, const void *fin, void *fout, const float *inlet_u, float *ux, float *uy,
                      float *rho)
{
    if (P.base) return launch_pair_fmt<FMT_R>(P, p->model, p->bc, c.chunk, st, mask, fin, fout, inlet_u, ux, uy, rho);
    return launch_any<FMT>(P, p, k, c, parity, st, mask, fin, fout, inlet_u, ux, uy, rho);
}

// One single step in place (the single-step kernel of lbm_step_aa / lbm_step_aa_h16)
static int single_step(const lbm_params *p, const uint8_t *mask, void *f, int fmt, const float *inlet_u, int parity,
                       float *ux, float *uy, float *rho, lbm_stream_t stream)
{
    return fmt == LBM_FMT_F32 ? lbm_step_aa(p, mask, static_cast<float *>(f), inlet_u, parity, ux, uy, rho, stream)
                              : lbm_step_aa_h16(p, mask, static_cast<uint16_t *>(f), fmt, inlet_u, parity, ux, uy, rho,
                                                stream);
}

// k steps that no single tile launch covers, without a configuration: from an even state one tile launch f_now ->
// f_out (the four-step kernel for FP32 populations with k >= 4, else the two-step one), from an odd state a device
// copy of f_now to f_out; then the remaining steps in place on f_out with the single-step kernel, the last one with
// the macros.  When the caller lets f_now be overwritten (f_scratch = f_now, the launches of lbm_step_fused_n), an
// odd state with k >= 3 takes its first step in place on f_now and the tile launch after it instead of the copy.
// Each piece is bit-identical to its single steps, so the whole is.  At 4096 x 1024 (FP32) single steps take 45 us
// and the copy 67 us, against 64-100 us per step for the strip sweep at k = 3..8 and 228 us at k = 1; from an even
// state k = 6 runs at 31.5-33.7 us per step, k = 8 at 35.8.
static int run_composed(const KPar &P, const lbm_params *p, const uint8_t *mask, const void *f_now, void *f_scratch,
                        void *f_out, int fmt, const float *inlet_u, int parity, int k, float *ux, float *uy,
                        float *rho, lbm_stream_t stream)
{
    const hipStream_t st = to_hip(stream);
    const bool f32 = fmt == LBM_FMT_F32;
    int done = 0, rc = LBM_OK;
    if (parity == 1 && f_scratch == f_now && f_scratch && k >= 3) {
        rc = single_step(p, mask, f_scratch, fmt, inlet_u, 1, nullptr, nullptr, nullptr, stream);
        if (rc) return rc;
        done = 1;
    }
    if (((parity + done) & 1) == 0 && k - done >= 2) {
        const int kt = f32 && k - done >= 4 ? 4 : 2;
        const lbm_fused_cfg t{0, 12};
        const bool last = done + kt == k;
        float *mx = last ? ux : nullptr, *my = last ? uy : nullptr, *mr = last ? rho : nullptr;
        if (f32)
            rc = launch_any<FMT_F32>(P, p, kt, t, 0, st, mask, f_now, f_out, inlet_u, mx, my, mr);
        else if (fmt == LBM_FMT_F16)
            rc = launch_h16<FMT_F16, FMT_F16R>(P, p, kt, t, 0, st, mask, f_now, f_out, inlet_u, mx, my, mr);
        else
            rc = launch_h16<FMT_BF16, FMT_BF16R>(P, p, kt, t, 0, st, mask, f_now, f_out, inlet_u, mx, my, mr);
        if (rc) return rc;
        rc = check_launch("lbm_step_fused");
        done += kt;
    } else {
        const size_t n = (size_t)(f32 ? 4 : 2) * 9 * (size_t)P.plane;
        const hipError_t e = hipMemcpyAsync(f_out, f_now, n, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return fail((int)e, "lbm_step_fused: state copy: %s", hipGetErrorString(e));
    }
    for (; done < k && rc == LBM_OK; ++done) {
        const bool last = done == k - 1;
        rc = single_step(p, mask, f_out, fmt, inlet_u, (parity + done) & 1, last ? ux : nullptr, last ? uy : nullptr,
                         last ? rho : nullptr, stream);
    }
    return rc;
}

// lbm_step_fused_cfg / lbm_step_fused_h16 (fmt = LBM_FMT_F32, F16 or BF16); f_scratch = f_now when f_now may be
// overwritten (run_composed), else NULL
static int fused_any(const lbm_params *p, const uint8_t *mask, const void *f_now, void *f_scratch, void *f_out,
                     int fmt, const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                     float *rho, lbm_stream_t stream)
{
    KPar P;
    lbm_fused_cfg c;
    bool composed;
    const bool f32 = fmt == LBM_FMT_F32;
    int rc = fused_common(p, mask, f_now, f_out, f32 ? 4 : 2, inlet_u, parity, k, cfg, ux, uy, rho, P, c, composed);
    if (rc) return rc;
    if (composed) return run_composed(P, p, mask, f_now, f_scratch, f_out, fmt, inlet_u, parity, k, ux, uy, rho, stream);
    const hipStream_t st = to_hip(stream);
    if (f32) {
        rc = launch_any<FMT_F32>(P, p, k, c, parity, st, mask, f_now, f_out, inlet_u, ux, uy, rho);
        if (rc) return rc;
        return check_launch("lbm_step_fused");
    }
    if (c.strip != 0 && c.strip != 128) // the 16-bit strip-sweep kernels are built for 128-column strips only
        return fail(LBM_ERR_UNSUPPORTED, "lbm_step_fused_h16: 16-bit storage takes strip 0 or 128, not %d", c.strip);
    // row-based storage runs the two-step tile kernel (checked in fused_common)
    if (fmt == LBM_FMT_F16)
        rc = launch_h16<FMT_F16, FMT_F16R>(P, p, k, c, parity, st, mask, f_now, f_out, inlet_u, ux, uy, rho);
    else
        rc = launch_h16<FMT_BF16, FMT_BF16R>(P, p, k, c, parity, st, mask, f_now, f_out, inlet_u, ux, uy, rho);
    if (rc) return rc;
    return check_launch("lbm_step_fused_h16");
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_fused_max_k(void) { return FUSED_MAXK; }

int lbm_fused_default_cfg(const lbm_params *p, int k, lbm_fused_cfg *cfg)
{
    if (!p || !cfg) return fail(LBM_ERR_ARG, "lbm_fused_default_cfg: NULL pointer");
    if (k < 1 || k > FUSED_MAXK) return fail(LBM_ERR_ARG, "lbm_fused_default_cfg: k=%d outside [1, %d]", k, FUSED_MAXK);
    if (p->W < 3 || p->H < 3) return fail(LBM_ERR_SIZE, "lbm_fused_default_cfg: lattice %dx%d too small", p->W, p->H);
    default_cfg(p->W, p->H, k, cfg);
    return LBM_OK;
}

int lbm_step_fused_cfg(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out,
                       const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                       float *rho, lbm_stream_t stream)
{
    return fused_any(p, mask, f_now, nullptr, f_out, LBM_FMT_F32, inlet_u, parity, k, cfg, ux, uy, rho, stream);
}

int lbm_step_fused(const lbm_params *p, const uint8_t *mask, const float *f_now, float *f_out, const float *inlet_u,
                   int parity, int k, float *ux, float *uy, float *rho, lbm_stream_t stream)
{
    return fused_any(p, mask, f_now, nullptr, f_out, LBM_FMT_F32, inlet_u, parity, k, nullptr, ux, uy, rho, stream);
}

int lbm_step_fused_h16(const lbm_params *p, const uint8_t *mask, const uint16_t *h_now, uint16_t *h_out, int fmt,
                       const float *inlet_u, int parity, int k, const lbm_fused_cfg *cfg, float *ux, float *uy,
                       float *rho, lbm_stream_t stream)
{
    if (fmt != LBM_FMT_F16 && fmt != LBM_FMT_BF16) return fail(LBM_ERR_ARG, "lbm_step_fused_h16: bad format %d", fmt);
    return fused_any(p, mask, h_now, nullptr, h_out, fmt, inlet_u, parity, k, cfg, ux, uy, rho, stream);
}

// the buffer a launch reads holds the previous launch's output (or the caller's state): scratch for run_composed
int lbm_step_fused_n(const lbm_params *p, const uint8_t *mask, float *f_a, float *f_b, const float *inlet_u,
                     int parity, int k, int nlaunch, const lbm_fused_cfg *cfg, lbm_stream_t stream)
{
    if (nlaunch < 0) return fail(LBM_ERR_ARG, "lbm_step_fused_n: nlaunch must be >= 0");
    for (int i = 0; i < nlaunch; ++i) {
        float *src = i & 1 ? f_b : f_a, *dst = i & 1 ? f_a : f_b;
        int rc = fused_any(p, mask, src, src, dst, LBM_FMT_F32, inlet_u, (parity + i * k) & 1, k, cfg, nullptr,
                           nullptr, nullptr, stream);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm_step_fused_h16_n(const lbm_params *p, const uint8_t *mask, uint16_t *h_a, uint16_t *h_b, int fmt,
                         const float *inlet_u, int parity, int k, int nlaunch, const lbm_fused_cfg *cfg,
                         lbm_stream_t stream)
{
    if (nlaunch < 0) return fail(LBM_ERR_ARG, "lbm_step_fused_h16_n: nlaunch must be >= 0");
    if (fmt != LBM_FMT_F16 && fmt != LBM_FMT_BF16) return fail(LBM_ERR_ARG, "lbm_step_fused_h16_n: bad format %d", fmt);
    for (int i = 0; i < nlaunch; ++i) {
        uint16_t *src = i & 1 ? h_b : h_a, *dst = i & 1 ? h_a : h_b;
        int rc = fused_any(p, mask, src, src, dst, fmt, inlet_u, (parity + i * k) & 1, k, cfg, nullptr, nullptr,
                           nullptr, stream);
        if (rc) return rc;
    }
    return LBM_OK;
}

} // extern "C"
