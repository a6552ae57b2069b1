// Host-side helpers shared by the liblbm translation units: argument checks, error reporting,
// parameter derivation and the launch-shape constants of the single-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/lbm.h"
#include "lbm_device.h"

namespace lbm {

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);
int check_launch(const char *what);

// Validates the parameter block and derives the kernel constants (FP32, host side).  pop_bytes is the
// size of one stored population (4 FP32, 2 for 16-bit deviations) for kernels that address the populations
// through buffer resources, whose record size and byte offsets are unsigned 32-bit: 9*W*H*pop_bytes must
// stay below 2^32.  pop_bytes = 0 checks only the element limit 9*W*H < 2^31 (pointer-indexed kernels).
int make_kpar(const lbm_params *p, KPar *k, unsigned pop_bytes);

inline hipStream_t to_hip(lbm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Single-step tiles: 256-thread blocks of 128 x 2 cells (each wave64 covers half a 512-byte row segment
// per plane), measured fastest among 32x8 .. 256x2 on MI355X.
constexpr int FUSED_MAXK = 8; // largest k of the temporally blocked kernel

constexpr int TILE_X = 128;
constexpr int TILE_Y = 2;

// Blocks are dealt round-robin to the 8 XCDs in dispatch order (block b runs on XCD b % 8).  Renumber
// them so that each XCD walks a contiguous band of rows: the odd AA step reads and writes rows y-1..y+1,
// and neighbouring rows then hit the same XCD's L2 instead of being fetched by two XCDs.
__device__ __forceinline__ void xcd_tile(int &bx, int &by)
{
    const int nbx = gridDim.x, nb = gridDim.x * gridDim.y;
    const int b = (int)blockIdx.y * nbx + (int)blockIdx.x;
    const int per = nb / 8;
    const int t = b < per * 8 ? (b % 8) * per + b / 8 : b;
    bx = t % nbx;
    by = t / nbx;
}

// The single step's renumbering: rows of tiles dealt round-robin to the XCDs (tile row r on XCD r % 8), each
// XCD walking its rows left to right.  Consecutive tiles of a row -- whose shifted odd-step writes share the
// 128-byte lines at their common edge -- stay on one L2, while the eight XCDs together sweep one compact band
// of rows instead of eight bands an eighth of the lattice apart.  Beyond the 256 MB MALL (4096 x 2048 to
// 8192 x 4096 FP32, 302 MB to 2.4 GB) this runs the single step 5-10 % faster than the banded xcd_tile; inside
// it (226 MB and less) 1-2 % slower.  Tile rows past the last whole round of 8 keep dispatch order.
__device__ __forceinline__ void xcd_row_tile(int &bx, int &by)
{
    const int nbx = gridDim.x;
    const int b = (int)blockIdx.y * nbx + (int)blockIdx.x;
    const int full = ((int)gridDim.y / 8) * 8 * nbx; // tiles of the whole rounds of 8 tile rows
    int t = b;
    if (b < full) {
        const int k = b % 8, j = b / 8; // XCD, position in that XCD's sequence
        t = ((j / nbx) * 8 + k) * nbx + j % nbx;
    }
    bx = t % nbx;
    by = t / nbx;
}

} // namespace lbm
