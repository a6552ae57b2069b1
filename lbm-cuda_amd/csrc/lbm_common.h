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

// Validates the parameter block and derives the kernel constants (FP32, host side).
int make_kpar(const lbm_params *p, KPar *k);

inline hipStream_t to_hip(lbm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Single-step tiles: one wave64 per row segment, four rows per 256-thread block.
constexpr int TILE_X = 64;
constexpr int TILE_Y = 4;

} // namespace lbm
