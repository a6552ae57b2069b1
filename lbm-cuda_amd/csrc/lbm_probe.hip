// Bandwidth ceiling of the in-place single step's memory pattern, with no collision: the single step's
// 128 x 2-cell tiles (WIDTH x 128 x 2 for wider accesses) in its XCD order (xcd_tile, or xcd_row_tile for states
// beyond the 256 MB MALL).  Every thread reads WIDTH consecutive FP32 values of each of the 9 planes through one
// buffer resource (as the AA step does: all nine loads, then all nine stores), multiplies them by s and stores
// them back to the same addresses.  SHIFT moves plane q's window by the lattice link c_q (x + y * W, as the odd
// step's neighbour accesses; windows past either end of the plane stay unshifted), so that unaligned wide
// accesses are measured too.  No reference counterpart: it is the measured ceiling the single step is compared
// with.
#include "lbm_common.h"

namespace lbm {

template <int WIDTH> struct Vec;
template <> struct Vec<1> {
    using T = uint32_t;
    __device__ static T ld(rsrc_t r, unsigned v, unsigned s) { return __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0); }
    __device__ static void st(rsrc_t r, unsigned v, unsigned s, T x) { __builtin_amdgcn_raw_buffer_store_b32(x, r, v, s, 0); }
    __device__ static uint32_t &at(T &x, int) { return x; }
};
template <> struct Vec<2> {
    using T = __attribute__((ext_vector_type(2))) uint32_t;
    __device__ static T ld(rsrc_t r, unsigned v, unsigned s) { return __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0); }
    __device__ static void st(rsrc_t r, unsigned v, unsigned s, T x) { __builtin_amdgcn_raw_buffer_store_b64(x, r, v, s, 0); }
    __device__ static uint32_t &at(T &x, int j) { return reinterpret_cast<uint32_t *>(&x)[j]; }
};
template <> struct Vec<4> {
    using T = __attribute__((ext_vector_type(4))) uint32_t;
    __device__ static T ld(rsrc_t r, unsigned v, unsigned s) { return __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0); }
    __device__ static void st(rsrc_t r, unsigned v, unsigned s, T x) { __builtin_amdgcn_raw_buffer_store_b128(x, r, v, s, 0); }
    __device__ static uint32_t &at(T &x, int j) { return reinterpret_cast<uint32_t *>(&x)[j]; }
};

template <int WIDTH, bool SHIFT>
__global__ __launch_bounds__(TILE_X *TILE_Y) void k_rmw_probe(float *buf, int W, int H, unsigned planeB, float s, int rowmap)
{
    using V = Vec<WIDTH>;
    const rsrc_t r = make_rsrc(buf, 9u * planeB);
    int bx, by;
    if (rowmap)
        xcd_row_tile(bx, by);
    else
        xcd_tile(bx, by);
    const int x = (bx * TILE_X + (int)threadIdx.x) * WIDTH, y = by * TILE_Y + (int)threadIdx.y;
    if (x >= W || y >= H) return;
    const long long cells = (long long)W * H, c0 = (long long)y * W + x;
    typename V::T v[9];
    unsigned off[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        long long c = c0;
        if constexpr (SHIFT) {
            c += cx(q) + (long long)cy(q) * W;
            if (c < 0 || c + WIDTH > cells) c = c0;
        }
        off[q] = (unsigned)c * 4u;
        v[q] = V::ld(r, off[q], (unsigned)q * planeB);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
#pragma unroll
        for (int j = 0; j < WIDTH; ++j)
            V::at(v[q], j) = __builtin_bit_cast(uint32_t, __builtin_bit_cast(float, V::at(v[q], j)) * s);
        V::st(r, off[q], (unsigned)q * planeB, v[q]);
    }
}

template <int WIDTH>
static void launch_probe(bool shift, float *buf, int W, int H, unsigned planeB, float s, hipStream_t st)
{
    const int rowmap = 9ull * planeB > (256ull << 20) ? 1 : 0; // as make_kpar decides for the single step
    const dim3 grid((W / WIDTH + TILE_X - 1) / TILE_X, (H + TILE_Y - 1) / TILE_Y), block(TILE_X, TILE_Y);
    if (shift)
        hipLaunchKernelGGL((k_rmw_probe<WIDTH, true>), grid, block, 0, st, buf, W, H, planeB, s, rowmap);
    else
        hipLaunchKernelGGL((k_rmw_probe<WIDTH, false>), grid, block, 0, st, buf, W, H, planeB, s, rowmap);
}

} // namespace lbm

using namespace lbm;

extern "C" {

int lbm_rmw_probe(float *buf, int W, int H, int width, int shift, float s, lbm_stream_t stream)
{
    if (!buf) return fail(LBM_ERR_ARG, "lbm_rmw_probe: NULL buffer");
    if (W < 1 || H < 1) return fail(LBM_ERR_ARG, "lbm_rmw_probe: W and H must be >= 1");
    if (width != 1 && width != 2 && width != 4) return fail(LBM_ERR_ARG, "lbm_rmw_probe: width must be 1, 2 or 4");
    if (W % width) return fail(LBM_ERR_ARG, "lbm_rmw_probe: W must be a multiple of width");
    if ((long long)W * H * 9 * 4 >= (1LL << 32))
        return fail(LBM_ERR_SIZE, "lbm_rmw_probe: 9 * W * H * 4 bytes must be < 2^32");
    hipStream_t st = to_hip(stream);
    const unsigned planeB = (unsigned)W * (unsigned)H * 4u;
    switch (width) {
    case 1: launch_probe<1>(shift != 0, buf, W, H, planeB, s, st); break;
    case 2: launch_probe<2>(shift != 0, buf, W, H, planeB, s, st); break;
    default: launch_probe<4>(shift != 0, buf, W, H, planeB, s, st); break;
    }
    return check_launch("lbm_rmw_probe");
}

} // extern "C"
