// The learned subgrid closure inside the collide-stream step (periodic lattices, FP32 state):
//   lbm_neq           m7neq / m8neq of the pre-collision state at every cell (lbm_core4.cu:166-173)
//   lbm_step_closure  one in-place AA step whose stress moments relax at 1 / (tau_0 + 3 nu_t), with nu_t
//                     from the MLP of lbm_mlp.h evaluated on the 3 x 3 stencil of (m7neq, m8neq) and clamped at
//                     0; cells whose features or output leave the trusted range fall back to Smagorinsky
//                     (lbm_core4.cu:171-177)
//
// No reference counterpart beyond the Smagorinsky fallback.  A step is two launches: k_neq writes the two
// non-equilibrium moments of every cell (the features must be those of the current state, which for an odd
// step only exist after the gather), then k_closure builds each cell's 18 features, runs the network on the
// matrix cores (one wave = 64 cells) and performs the cell's AA update with the external rate.
#include "lbm_aa_cell.h"
#include "lbm_mlp.h"

namespace lbm {

// m7neq and m8neq of the deviations h, in the operation order of collide_mrt (so the features equal the
// moments the collision relaxes).
__device__ __forceinline__ void mrt_neq(const KPar &P, const float h[9], float &n7, float &n8)
{
    const float s13 = h[1] + h[3], s24 = h[2] + h[4], s57 = h[5] + h[7], s68 = h[6] + h[8];
    const float h14 = s13 + s24, h58 = s57 + s68;
    const float a13 = h[1] - h[3], a24 = h[2] - h[4], a57 = h[5] - h[7], a86 = h[8] - h[6];
    const float sx = a57 + a86, sy = a57 - a86;
    const float drho = (h[0] + h14) + h58;
    const float rho = 1.0f + drho;
    const float jx = a13 + sx, jy = a24 + sy;
    const float inv_rho = recip_rn(rho);
    float ux = jx * inv_rho, uy = jy * inv_rho, drc = drho;
    if (P.clamp) {
        ux = clampf(ux, -0.4f, 0.4f);
        uy = clampf(uy, -0.4f, 0.4f);
        drc = clampf(drho, -1.0f, 0.5f);
    }
    const float rc = 1.0f + drc;
    const float m3 = rc * ux;
    const float ux2 = ux * ux, uy2 = uy * uy;
    n7 = (s13 - s24) - rc * (ux2 - uy2);
    n8 = (s57 - s68) - m3 * uy;
}

template <int PARITY>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_neq(KPar P, const uint8_t *__restrict__ mask,
                                                        const float *__restrict__ f, float *__restrict__ neq)
{
    const int x = blockIdx.x * TILE_X + threadIdx.x, y = blockIdx.y * TILE_Y + threadIdx.y;
    if (x >= P.W || y >= P.H) return;
    const int W = P.W;
    const long plane = P.plane;
    const long idx = (long)y * W + x;
    if (mask[idx]) {
        neq[idx] = 0.0f;
        neq[plane + idx] = 0.0f;
        return;
    }
    float h[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        if (PARITY == 0 || q == 0) {
            h[q] = f[q * plane + idx] - wq(q);
        } else { // incoming f_q = slot opp(q) of the periodic source x - c_q, or this cell's slot q (bounce-back)
            int sx = x - cx(q), sy = y - cy(q);
            sx = sx < 0 ? sx + W : (sx >= W ? sx - W : sx);
            sy = sy < 0 ? sy + P.H : (sy >= P.H ? sy - P.H : sy);
            const long s = (long)sy * W + sx;
            h[q] = (mask[s] ? f[q * plane + idx] : f[opp(q) * plane + s]) - wq(q);
        }
    }
    float n7, n8;
    mrt_neq(P, h, n7, n8);
    neq[idx] = n7;
    neq[plane + idx] = n8;
}

template <int PARITY>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_closure(KPar P, const uint8_t *__restrict__ mask,
                                                            float *__restrict__ f, const float *__restrict__ neq,
                                                            const uint16_t *__restrict__ pk,
                                                            const float *__restrict__ prm, lbm_closure cl,
                                                            float *__restrict__ nut, unsigned *__restrict__ nfall)
{
    using namespace mlp;
    __shared__ __align__(16) uint16_t wl[NPACK];
    __shared__ __align__(16) uint16_t tiles[4][WAVE_LDS];
    __shared__ float ys[4][64];
    const int t = threadIdx.y * TILE_X + threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    stage_weights(wl, pk, NPACK);
    const int x = blockIdx.x * TILE_X + threadIdx.x, y = blockIdx.y * TILE_Y + threadIdx.y;
    const int W = P.W, H = P.H;
    const bool in = x < W && y < H;
    const long plane = P.plane;
    // features: (m7neq, m8neq) at (x + dx, y + dy), dy, dx in -1..1, row-major, scaled
    uint16_t *xs = tiles[w], *h1s = xs + 64 * XS, *h2s = h1s + 64 * HS;
    bool trusted = in;
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        float a = 0.0f, b = 0.0f;
        if (in) {
            int nx = x + (s % 3) - 1, ny = y + (s / 3) - 1;
            nx = nx < 0 ? nx + W : (nx >= W ? nx - W : nx);
            ny = ny < 0 ? ny + H : (ny >= H ? ny - H : ny);
            const long j = (long)ny * W + nx;
            a = neq[j] * cl.in_scale[2 * s];
            b = neq[plane + j] * cl.in_scale[2 * s + 1];
        }
        trusted = trusted && fabsf(a) <= cl.feat_max && fabsf(b) <= cl.feat_max;
        xs[lane * XS + 2 * s] = mlp::f2bf(a);
        xs[lane * XS + 2 * s + 1] = mlp::f2bf(b);
    }
#pragma unroll
    for (int k = NIN; k < KIN; ++k) xs[lane * XS + k] = 0;
    __syncthreads();
    f4 z1[4][4], z2[4][4];
    wave_forward<false>(xs, h1s, h2s, wl, prm, ys[w], z1, z2);
    __syncthreads();
    // negative outputs (backscatter) clamp to no eddy viscosity; NaN, nu > nu_max or untrusted features
    // fall back to Smagorinsky
    const float raw = ys[w][lane] * cl.nu_scale;
    const bool ok = trusted && raw <= cl.nu_max; // false for NaN
    const float nu = raw > 0.0f ? raw : 0.0f;
    float w_ext = -1.0f;
    if (ok) w_ext = recip_rn(P.tau0 + 3.0f * nu);
    if (nfall) {
        const unsigned long long bad = __ballot(in && !ok && !mask[(long)y * W + x]);
        if (lane == 0 && bad) atomicAdd(nfall, (unsigned)__popcll(bad));
    }
    if (in && nut) nut[(long)y * W + x] = ok ? nu : -1.0f;
    aa_cell<MODEL_MRT, FMT_F32, PARITY, true, false, true>(P, mask, f, nullptr, nullptr, nullptr, nullptr, x, y,
                                                           w_ext);
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_neq(const lbm_params *p, const uint8_t *mask, const float *f, int parity, float *neq, lbm_stream_t stream)
{
    KPar P;
    int rc = make_kpar(p, &P, 4);
    if (rc) return rc;
    if (p->bc != LBM_BC_PERIODIC || p->model != LBM_MRT_LES || P.sums || P.fast)
        return fail(LBM_ERR_UNSUPPORTED, "lbm_neq: periodic MRT lattices with plain sums and exact arithmetic only");
    if (!mask || !f || !neq) return fail(LBM_ERR_ARG, "lbm_neq: NULL buffer");
    if (parity != 0 && parity != 1) return fail(LBM_ERR_ARG, "lbm_neq: parity must be 0 or 1");
    dim3 grid((P.W + TILE_X - 1) / TILE_X, (P.H + TILE_Y - 1) / TILE_Y);
    if (parity)
        hipLaunchKernelGGL(k_neq<1>, grid, dim3(TILE_X, TILE_Y), 0, to_hip(stream), P, mask, f, neq);
    else
        hipLaunchKernelGGL(k_neq<0>, grid, dim3(TILE_X, TILE_Y), 0, to_hip(stream), P, mask, f, neq);
    return check_launch("lbm_neq");
}

int lbm_step_closure(const lbm_params *p, const uint8_t *mask, float *f, int parity, const uint16_t *packed,
                     const float *params, const lbm_closure *cl, float *neq, float *nut, unsigned *fallbacks,
                     lbm_stream_t stream)
{
    if (!packed || !params || !cl) return fail(LBM_ERR_ARG, "lbm_step_closure: NULL model");
    if (!(cl->nu_max >= 0.0f) || !(cl->feat_max >= 0.0f))
        return fail(LBM_ERR_ARG, "lbm_step_closure: nu_max and feat_max must be >= 0");
    int rc = lbm_neq(p, mask, f, parity, neq, stream);
    if (rc) return rc;
    KPar P;
    rc = make_kpar(p, &P, 4);
    if (rc) return rc;
    if (P.fast) return fail(LBM_ERR_UNSUPPORTED, "lbm_step_closure: fast arithmetic is not built for the closure step");
    dim3 grid((P.W + TILE_X - 1) / TILE_X, (P.H + TILE_Y - 1) / TILE_Y);
    if (parity)
        hipLaunchKernelGGL(k_closure<1>, grid, dim3(TILE_X, TILE_Y), 0, to_hip(stream), P, mask, f, neq, packed,
                           params, *cl, nut, fallbacks);
    else
        hipLaunchKernelGGL(k_closure<0>, grid, dim3(TILE_X, TILE_Y), 0, to_hip(stream), P, mask, f, neq, packed,
                           params, *cl, nut, fallbacks);
    return check_launch("lbm_step_closure");
}

} // extern "C"
