// Training and inference kernels of the subgrid-closure MLP (layout and conventions in lbm_mlp.h):
//   lbm_mlp_pack        FP32 parameters -> packed BF16 weights
//   lbm_mlp_forward     y = net(x) for n samples
//   lbm_mlp_grad        gradient of L = sum_s wgt_s (y_s - t_s)^2 over n samples (no update)
//   lbm_mlp_train_step  the same gradient, then one Adam step on the FP32 parameters and a repack
//
// No reference counterpart (the reference has no learned closure).  Each wave takes 64 samples: forward on
// the matrix cores keeping the pre-activations in registers, then backward with the weight gradients as
// [K x 64] x [64 x N] products over the wave's samples (BF16 operands, FP32 accumulation).  Every wave writes
// its own partial gradient; one more kernel sums the partials in a fixed order (deterministic) and applies
// Adam.
#include "lbm_mlp.h"

namespace lbm {
namespace mlp {

constexpr int WAVES = 4; // per workgroup
constexpr int PART = NPARAM + 1; // one partial: gradient + loss

// Zero-padded BF16 pack of the FP32 parameters (one thread per packed value).
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ prm, uint16_t *__restrict__ pk)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= NPACK) return;
    float v = 0.0f;
    if (i < PK_W2T) {
        const int n = (i - PK_W1T) / KIN, k = (i - PK_W1T) % KIN;
        if (k < NIN) v = prm[OFF_W1 + k * NH + n];
    } else if (i < PK_W2N) {
        const int n = (i - PK_W2T) / NH, k = (i - PK_W2T) % NH;
        v = prm[OFF_W2 + k * NH + n];
    } else if (i < PK_W3T) {
        v = prm[OFF_W2 + (i - PK_W2N)];
    } else {
        const int n = (i - PK_W3T) / NH, k = (i - PK_W3T) % NH;
        if (n == 0) v = prm[OFF_W3 + k];
    }
    pk[i] = f2bf(v);
}

// Loads the wave's 64 samples (rows s0 .. s0+63 of x[n][18]) into xs as BF16; rows past n are zero.
__device__ __forceinline__ void load_inputs(uint16_t *xs, const float *__restrict__ x, long s0, int n)
{
    const int lane = threadIdx.x & 63;
    const long s = s0 + lane;
#pragma unroll
    for (int k = 0; k < KIN; ++k) xs[lane * XS + k] = (k < NIN && s < n) ? f2bf(x[s * NIN + k]) : (uint16_t)0;
}

__global__ __launch_bounds__(256) void k_forward(const uint16_t *__restrict__ pk, const float *__restrict__ prm,
                                                 const float *__restrict__ x, int n, float *__restrict__ y)
{
    __shared__ __align__(16) uint16_t wl[NPACK];
    __shared__ __align__(16) uint16_t tiles[WAVES][WAVE_LDS];
    __shared__ float ys[WAVES][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    stage_weights(wl, pk, NPACK);
    uint16_t *xs = tiles[w], *h1s = xs + 64 * XS, *h2s = h1s + 64 * HS;
    const long s0 = ((long)blockIdx.x * WAVES + w) * 64;
    load_inputs(xs, x, s0, n);
    __syncthreads();
    f4 z1[4][4], z2[4][4];
    wave_forward<false>(xs, h1s, h2s, wl, prm, ys[w], z1, z2);
    __syncthreads();
    if (s0 + lane < n) y[s0 + lane] = ys[w][lane];
}

// Fused forward + backward; writes one partial gradient per wave to part[wave][PART].
__global__ __launch_bounds__(256) void k_grad(const uint16_t *__restrict__ pk, const float *__restrict__ prm,
                                              const float *__restrict__ x, const float *__restrict__ t,
                                              const float *__restrict__ wgt, int n, float *__restrict__ part)
{
    __shared__ __align__(16) uint16_t wl[NPACK];
    __shared__ __align__(16) uint16_t tiles[WAVES][WAVE_LDS];
    __shared__ float ys[WAVES][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    stage_weights(wl, pk, NPACK);
    uint16_t *xs = tiles[w], *h1s = xs + 64 * XS, *h2s = h1s + 64 * HS;
    const long wave = (long)blockIdx.x * WAVES + w;
    const long s0 = wave * 64;
    float *out = part + wave * PART;
    load_inputs(xs, x, s0, n);
    __syncthreads();
    f4 z1[4][4], z2[4][4];
    wave_forward<true>(xs, h1s, h2s, wl, prm, ys[w], z1, z2);
    __syncthreads();

    // loss and dL/dy
    const long s = s0 + lane;
    const float e = s < n ? ys[w][lane] - t[s] : 0.0f;
    const float ws = s < n ? wgt[s] : 0.0f;
    const float dy = 2.0f * ws * e;
    float loss = ws * e * e;
    float db3 = dy;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        loss += __shfl_xor(loss, o);
        db3 += __shfl_xor(db3, o);
    }
    __syncthreads(); // every lane has read ys before it is reused
    ys[w][lane] = dy;
    __syncthreads();
    // dW3[k] = sum_s h2[s][k] dy[s]   (lane k)
    {
        float acc = 0.0f;
        for (int i = 0; i < 64; ++i) acc = fmaf(bf2f(h2s[i * HS + lane]), ys[w][i], acc);
        out[OFF_W3 + lane] = acc;
    }
    if (lane == 0) {
        out[OFF_B3] = db3;
        out[NPARAM] = loss;
    }
    __syncthreads(); // h2s is overwritten by dZ2 below
    // dZ2 = dy W3^T * gelu'(z2)  (D layout), db2
    float db[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float w3 = bf2f(wl[PK_W3T + 16 * nt + r]);
        float acc = 0.0f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = 16 * mt + 4 * g + j;
                const float dz = ys[w][i] * w3 * gelu_d(z2[mt][nt][j]);
                acc += dz;
                h2s[i * HS + 16 * nt + r] = f2bf(dz);
            }
        }
        db[nt] = acc;
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        db[nt] += __shfl_xor(db[nt], 16);
        db[nt] += __shfl_xor(db[nt], 32);
        if (g == 0) out[OFF_B2 + 16 * nt + r] = db[nt];
    }
    __syncthreads();
    // dW2[k][n] = sum_s h1[s][k] dZ2[s][n]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mma(ld8col(h1s, 32 * ks + 8 * g, HS, 16 * mt + r), ld8col(h2s, 32 * ks + 8 * g, HS, 16 * nt + r),
                          acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) out[OFF_W2 + (16 * mt + 4 * g + j) * NH + 16 * nt + r] = acc[j];
        }
    }
    // dZ1 = (dZ2 W2^T) * gelu'(z1), db1
    f4 dz1[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        b8 b[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) b[ks] = ld8(wl + PK_W2N + (16 * nt + r) * NH + 32 * ks + 8 * g);
        float acc_b = 0.0f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) acc = mma(ld8(h2s + (16 * mt + r) * HS + 32 * ks + 8 * g), b[ks], acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dz1[mt][nt][j] = acc[j] * gelu_d(z1[mt][nt][j]);
                acc_b += dz1[mt][nt][j];
            }
        }
        db[nt] = acc_b;
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        db[nt] += __shfl_xor(db[nt], 16);
        db[nt] += __shfl_xor(db[nt], 32);
        if (g == 0) out[OFF_B1 + 16 * nt + r] = db[nt];
    }
    __syncthreads(); // h1s is overwritten by dZ1
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) h1s[(16 * mt + 4 * g + j) * HS + 16 * nt + r] = f2bf(dz1[mt][nt][j]);
    __syncthreads();
    // dW1[k][n] = sum_s x[s][k] dZ1[s][n]   (k < 18 kept)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                acc = mma(ld8col(xs, 32 * ks + 8 * g, XS, 16 * mt + r), ld8col(h1s, 32 * ks + 8 * g, HS, 16 * nt + r),
                          acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * mt + 4 * g + j;
                if (k < NIN) out[OFF_W1 + k * NH + 16 * nt + r] = acc[j];
            }
        }
    }
}

// Sums the partials in wave order; optionally stores the gradient and the loss, and applies Adam
// (m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, p -= lr (m / bc1) / (sqrt(v / bc2) + eps)) and repacks.
__global__ __launch_bounds__(256) void k_reduce_adam(const float *__restrict__ part, int nparts,
                                                     float *__restrict__ grad, float *__restrict__ loss,
                                                     float *__restrict__ prm, float *__restrict__ m,
                                                     float *__restrict__ v, lbm_adam hp, uint16_t *__restrict__ pk)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > NPARAM) return;
    float acc = 0.0f;
    for (int i = 0; i < nparts; ++i) acc += part[(long)i * PART + p];
    if (p == NPARAM) {
        if (loss) *loss = acc;
        return;
    }
    if (grad) grad[p] = acc;
    if (!prm) return;
    const float mm = hp.beta1 * m[p] + (1.0f - hp.beta1) * acc;
    const float vv = hp.beta2 * v[p] + (1.0f - hp.beta2) * (acc * acc);
    m[p] = mm;
    v[p] = vv;
    const float np = prm[p] - hp.lr * (mm / hp.bc1) / (sqrtf(vv / hp.bc2) + hp.eps);
    prm[p] = np;
    const uint16_t b = f2bf(np);
    if (p < OFF_B1) {
        const int k = p / NH, n = p % NH;
        pk[PK_W1T + n * KIN + k] = b;
    } else if (p >= OFF_W2 && p < OFF_B2) {
        const int k = (p - OFF_W2) / NH, n = (p - OFF_W2) % NH;
        pk[PK_W2T + n * NH + k] = b;
        pk[PK_W2N + k * NH + n] = b;
    } else if (p >= OFF_W3 && p < OFF_B3) {
        pk[PK_W3T + (p - OFF_W3)] = b;
    }
}

static int nblocks_for(int n) { return (n + 64 * WAVES - 1) / (64 * WAVES); }

} // namespace mlp
} // namespace lbm

using namespace lbm;
using namespace lbm::mlp;

extern "C" {

int lbm_mlp_nparam(void) { return NPARAM; }
int lbm_mlp_npack(void) { return NPACK; }

int lbm_mlp_pack(const float *params, uint16_t *packed, lbm_stream_t stream)
{
    if (!params || !packed) return fail(LBM_ERR_ARG, "lbm_mlp_pack: NULL buffer");
    hipLaunchKernelGGL(k_pack, dim3((NPACK + 255) / 256), dim3(256), 0, to_hip(stream), params, packed);
    return check_launch("lbm_mlp_pack");
}

int lbm_mlp_forward(const uint16_t *packed, const float *params, const float *x, int n, float *y,
                    lbm_stream_t stream)
{
    if (!packed || !params || !x || !y) return fail(LBM_ERR_ARG, "lbm_mlp_forward: NULL buffer");
    if (n < 0) return fail(LBM_ERR_ARG, "lbm_mlp_forward: n < 0");
    if (n == 0) return LBM_OK;
    hipLaunchKernelGGL(k_forward, dim3(nblocks_for(n)), dim3(64 * WAVES), 0, to_hip(stream), packed, params, x, n, y);
    return check_launch("lbm_mlp_forward");
}

size_t lbm_mlp_work_size(int n)
{
    if (n <= 0) return 0;
    return (size_t)nblocks_for(n) * WAVES * PART * sizeof(float);
}

static int mlp_grad_common(const char *who, const uint16_t *packed, const float *params, const float *x,
                           const float *t, const float *wgt, int n, float *work, hipStream_t s)
{
    if (!packed || !params || !x || !t || !wgt || !work) return fail(LBM_ERR_ARG, "%s: NULL buffer", who);
    if (n <= 0) return fail(LBM_ERR_ARG, "%s: n must be positive", who);
    hipLaunchKernelGGL(k_grad, dim3(nblocks_for(n)), dim3(64 * WAVES), 0, s, packed, params, x, t, wgt, n, work);
    return LBM_OK;
}

int lbm_mlp_grad(const uint16_t *packed, const float *params, const float *x, const float *t, const float *wgt,
                 int n, float *work, float *grad, float *loss, lbm_stream_t stream)
{
    hipStream_t s = to_hip(stream);
    int rc = mlp_grad_common("lbm_mlp_grad", packed, params, x, t, wgt, n, work, s);
    if (rc) return rc;
    if (!grad) return fail(LBM_ERR_ARG, "lbm_mlp_grad: NULL gradient");
    lbm_adam none = {};
    hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, work,
                       nblocks_for(n) * WAVES, grad, loss, (float *)nullptr, (float *)nullptr, (float *)nullptr, none,
                       (uint16_t *)nullptr);
    return check_launch("lbm_mlp_grad");
}

int lbm_mlp_train_step(uint16_t *packed, float *params, float *m, float *v, const float *x, const float *t,
                       const float *wgt, int n, const lbm_adam *adam, float *work, float *grad, float *loss,
                       lbm_stream_t stream)
{
    if (!adam || !m || !v) return fail(LBM_ERR_ARG, "lbm_mlp_train_step: NULL optimiser state");
    if (!(adam->bc1 > 0.0f) || !(adam->bc2 > 0.0f))
        return fail(LBM_ERR_ARG, "lbm_mlp_train_step: bias corrections must be positive");
    hipStream_t s = to_hip(stream);
    int rc = mlp_grad_common("lbm_mlp_train_step", packed, params, x, t, wgt, n, work, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, work,
                       nblocks_for(n) * WAVES, grad, loss, params, m, v, *adam, packed);
    return check_launch("lbm_mlp_train_step");
}

} // extern "C"
