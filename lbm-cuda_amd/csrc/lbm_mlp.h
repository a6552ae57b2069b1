// The subgrid-closure MLP 18 -> 64 -> 64 -> 1 (GELU) on gfx950 matrix cores.
//
// No reference counterpart: the reference's only closure is the Smagorinsky model of lbm_core4.cu:171-177,
// whose inputs m7neq / m8neq (lbm_core4.cu:172-173) are the network's features here, taken over a 3 x 3
// stencil.  One wave evaluates 64 samples with gfx950's v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation, 32 of K per instruction): a layer is a [64 x K] x [K x 64] product, tiled 4 x 4 over 16 x 16
// output tiles.
//
// MFMA operand layout (16 x 16 x 32, lane l, r = l % 16, g = l / 16; A and B items j = 0..7, D items j = 0..3):
//   A[i][k]: lane holds A[r][8g + j]      B[k][n]: lane holds B[8g + j][r]      D[i][n]: lane holds D[4g + j][r]
//
// Parameters, FP32, in this order (5441 values):
//   W1 [18][64] (input-major) | b1 [64] | W2 [64][64] | b2 [64] | W3 [64] | b3
// z1 = x W1 + b1, h1 = gelu(z1); z2 = h1 W2 + b2, h2 = gelu(z2); y = h2 . W3 + b3;
// gelu(z) = z (1 + erf(z / sqrt 2)) / 2.
// The packed BF16 copy the kernels read (11264 values, round to nearest even):
//   W1t [64][32] = W1^T (inputs 18..31 zero) | W2t [64][64] = W2^T | W2n [64][64] = W2 | W3t [16][64] (row 0 = W3)
#pragma once

#include "lbm_common.h"

namespace lbm {
namespace mlp {

constexpr int NIN = 18, NH = 64, KIN = 32;
constexpr int OFF_W1 = 0, OFF_B1 = NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH,
              OFF_B3 = OFF_W3 + NH, NPARAM = OFF_B3 + 1;
constexpr int PK_W1T = 0, PK_W2T = PK_W1T + NH * KIN, PK_W2N = PK_W2T + NH * NH, PK_W3T = PK_W2N + NH * NH,
              NPACK = PK_W3T + 16 * NH;
static_assert(NPARAM == 5441 && NPACK == 11264, "parameter layout");

// LDS row strides (elements) of a wave's sample tiles; 8-byte aligned rows, staggered banks
constexpr int XS = 40; // inputs [64][XS], columns 0..31 used
constexpr int HS = 72; // activations [64][HS], columns 0..63 used
constexpr int WAVE_LDS = 64 * XS + 2 * 64 * HS; // bf16 elements per wave

typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even (inputs are finite)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

__device__ __forceinline__ f4 mma(b8 a, b8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// eight consecutive BF16 values (16-byte aligned: every row stride and offset used is a multiple of 8 elements)
__device__ __forceinline__ b8 ld8(const uint16_t *p) { return *reinterpret_cast<const b8 *>(p); }

// eight values of one column at rows row0 .. row0 + 7
__device__ __forceinline__ b8 ld8col(const uint16_t *p, int row0, int stride, int col)
{
    b8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = __builtin_bit_cast(__bf16, p[(row0 + j) * stride + col]);
    return v;
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Forward pass of one wave's 64 samples.  xs: [64][XS] BF16 inputs; h1s, h2s: [64][HS] activation tiles;
// wp: packed weights; prm: FP32 parameters (biases).  y[s] of sample s lands in ys[s] (written by the lanes
// with r == 0).  KEEP: the pre-activations stay in z1 / z2 (D layout) for the backward pass.  The caller
// synchronises between writing xs and calling, and before reading ys.
template <bool KEEP>
__device__ __forceinline__ void wave_forward(const uint16_t *xs, uint16_t *h1s, uint16_t *h2s, const uint16_t *wp,
                                             const float *__restrict__ prm, float *ys, f4 (&z1)[4][4],
                                             f4 (&z2)[4][4])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float bias = prm[OFF_B1 + 16 * nt + r];
        const b8 b0 = ld8(wp + PK_W1T + (16 * nt + r) * KIN + 8 * g);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma(ld8(xs + (16 * mt + r) * XS + 8 * g), b0, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float z = acc[j] + bias;
                if constexpr (KEEP) z1[mt][nt][j] = z;
                h1s[(16 * mt + 4 * g + j) * HS + 16 * nt + r] = f2bf(gelu(z));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float bias = prm[OFF_B2 + 16 * nt + r];
        b8 b[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) b[ks] = ld8(wp + PK_W2T + (16 * nt + r) * NH + 32 * ks + 8 * g);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) acc = mma(ld8(h1s + (16 * mt + r) * HS + 32 * ks + 8 * g), b[ks], acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float z = acc[j] + bias;
                if constexpr (KEEP) z2[mt][nt][j] = z;
                h2s[(16 * mt + 4 * g + j) * HS + 16 * nt + r] = f2bf(gelu(z));
            }
        }
    }
    __syncthreads();
    const float b3 = prm[OFF_B3];
    b8 b[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) b[ks] = ld8(wp + PK_W3T + r * NH + 32 * ks + 8 * g);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) acc = mma(ld8(h2s + (16 * mt + r) * HS + 32 * ks + 8 * g), b[ks], acc);
        if (r == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) ys[16 * mt + 4 * g + j] = acc[j] + b3;
        }
    }
}

// Copies the packed weights into LDS (the whole workgroup cooperates; `n` BF16 values from offset 0).
__device__ __forceinline__ void stage_weights(uint16_t *dst, const uint16_t *__restrict__ src, int n)
{
    const int nt = blockDim.x * blockDim.y;
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    for (int i = t; i < n / 2; i += nt) d[i] = s[i];
}

} // namespace mlp
} // namespace lbm
