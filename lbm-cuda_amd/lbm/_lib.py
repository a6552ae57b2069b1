"""ctypes binding of liblbm.so (declared in include/lbm.h).

The library is built in-tree by ``lbm.build`` (hipcc, gfx950).  There is no fallback: if the shared
library is missing or a symbol is absent, loading raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm.so")

MRT_LES, BGK = 0, 1
CHANNEL, PERIODIC = 0, 1
FMT_F32, FMT_F16, FMT_BF16 = 0, 1, 2

ABI_VERSION = 5
ROW_BASE_STRIDE = 16  # LBM_ROW_BASE_STRIDE


class LbmBaseWindow(ctypes.Structure):
    """Mirror of ``lbm_base_window`` (include/lbm.h)."""

    _fields_ = [("x0", ctypes.c_int32), ("y0", ctypes.c_int32), ("w", ctypes.c_int32), ("h", ctypes.c_int32),
                ("vel", ctypes.c_void_p)]


class LbmParams(ctypes.Structure):
    """Mirror of ``lbm_params`` (include/lbm.h)."""

    _fields_ = [
        ("W", ctypes.c_int32),
        ("H", ctypes.c_int32),
        ("model", ctypes.c_int32),
        ("bc", ctypes.c_int32),
        ("clamp", ctypes.c_int32),
        ("tau_inv", ctypes.c_float),
        ("cs", ctypes.c_float),
        ("s_e", ctypes.c_float),
        ("s_eps", ctypes.c_float),
        ("s_q", ctypes.c_float),
        ("sums", ctypes.c_int32),
        ("fast", ctypes.c_int32),
        ("row_base", ctypes.c_void_p),
        ("window", ctypes.POINTER(LbmBaseWindow)),
    ]


class LbmFusedCfg(ctypes.Structure):
    """Mirror of ``lbm_fused_cfg`` (include/lbm.h)."""

    _fields_ = [("strip", ctypes.c_int32), ("chunk", ctypes.c_int32)]


class LbmAdam(ctypes.Structure):
    """Mirror of ``lbm_adam`` (include/lbm.h)."""

    _fields_ = [(n, ctypes.c_float) for n in ("lr", "beta1", "beta2", "eps", "bc1", "bc2")]


class LbmClosure(ctypes.Structure):
    """Mirror of ``lbm_closure`` (include/lbm.h)."""

    _fields_ = [("in_scale", ctypes.c_float * 18), ("nu_scale", ctypes.c_float), ("nu_max", ctypes.c_float),
                ("feat_max", ctypes.c_float)]


class LbmError(RuntimeError):
    pass


_vp = ctypes.c_void_p
_P = ctypes.POINTER(LbmParams)
_i = ctypes.c_int
_f = ctypes.c_float

# name -> (restype, argtypes); every symbol declared in include/lbm.h
SIGNATURES = {
    "lbm_abi_version": (_i, []),
    "lbm_last_error": (ctypes.c_char_p, []),
    "lbm_sync": (_i, [_vp]),
    "lbm_step_aa": (_i, [_P, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "lbm_step_aa_n": (_i, [_P, _vp, _vp, _vp, _i, _i, _vp]),
    "lbm_step_pull": (_i, [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lbm_step_fused": (_i, [_P, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp]),
    "lbm_fused_max_k": (_i, []),
    "lbm_fused_default_cfg": (_i, [_P, _i, _vp]),
    "lbm_step_fused_cfg": (_i, [_P, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "lbm_macro": (_i, [_P, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "lbm_step_aa_h16": (_i, [_P, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp]),
    "lbm_step_aa_h16_n": (_i, [_P, _vp, _vp, _i, _vp, _i, _i, _vp]),
    "lbm_step_fused_h16": (_i, [_P, _vp, _vp, _vp, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "lbm_macro_h16": (_i, [_P, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp]),
    "lbm_convert": (_i, [_P, _vp, _i, _vp, _i, _vp]),
    "lbm_convert_state": (_i, [_P, _vp, _i, _vp, _i, _i, _vp]),
    "lbm_row_base": (_i, [_i, _vp, _vp, _vp]),
    "lbm_rebase_window": (_i, [_P, _vp, _i, _i, _vp, _vp]),
    "lbm_visualize": (_i, [_P, _vp, _vp, _vp, _vp, _f, _vp]),
    "lbm_diagnostics_work_size": (ctypes.c_size_t, [_P]),
    "lbm_diagnostics": (_i, [_P, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "lbm_diagnostics_scratch_bytes": (ctypes.c_size_t, []),
    "lbm_diagnostics_noalloc": (_i, [_P, _vp, _vp, _vp, _i, _vp, _vp]),
    "lbm_ensemble_create": (_i, [_vp, _i, _i, _vp]),
    "lbm_ensemble_info": (_i, [_vp, _vp, _vp, _vp]),
    "lbm_ensemble_step": (_i, [_vp, _i, _i, _vp]),
    "lbm_ensemble_step_n": (_i, [_vp, _i, _i, ctypes.c_longlong, _i, _vp]),
    "lbm_ensemble_destroy": (_i, [_vp]),
    "lbm_ensemble_pair_n": (_i, [_vp, _i, ctypes.c_longlong, _i, _i, _vp]),
    "lbm_ensemble_side": (_i, [_vp]),
    "lbm_ensemble_pair_info": (_i, [_vp, _i, _vp, _vp]),
    "lbm_ensemble_run": (_i, [_vp, _i, _i, _i, _i, _i, _i, ctypes.c_longlong, _i, _vp]),
    "lbm_ensemble_run_info": (_i, [_vp, _i, _i, _vp, _vp]),
    "lbm_ensemble_member_side": (_i, [_vp, _i]),
    "lbm_selftest": (_i, [_i, _vp, _vp, _vp]),
    "lbm_step_fused_n": (_i, [_P, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp]),
    "lbm_step_fused_h16_n": (_i, [_P, _vp, _vp, _vp, _i, _vp, _i, _i, _i, _vp, _vp]),
    "lbm_selftest_threads": (_i, []),
    "lbm_rmw_probe": (_i, [_vp, _i, _i, _i, _i, ctypes.c_float, _vp]),
    "lbm_graph_begin": (_i, [_vp]),
    "lbm_graph_end": (_i, [_vp, _vp]),
    "lbm_graph_launch": (_i, [_vp, _vp]),
    "lbm_graph_destroy": (_i, [_vp]),
    "lbm_mlp_nparam": (_i, []),
    "lbm_mlp_npack": (_i, []),
    "lbm_mlp_pack": (_i, [_vp, _vp, _vp]),
    "lbm_mlp_forward": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "lbm_mlp_work_size": (ctypes.c_size_t, [_i]),
    "lbm_mlp_grad": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "lbm_mlp_train_step": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "lbm_neq": (_i, [_P, _vp, _vp, _i, _vp, _vp]),
    "lbm_step_closure": (_i, [_P, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load liblbm.so and bind every entry point (raises if anything is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LbmError(f"liblbm.so not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)  # AttributeError if the symbol is absent
        fn.restype = res
        fn.argtypes = args
    v = L.lbm_abi_version()
    if v != ABI_VERSION:
        raise LbmError(f"liblbm ABI version {v}, expected {ABI_VERSION}")
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        msg = load().lbm_last_error().decode(errors="replace")
        raise LbmError(f"{what} failed ({rc}): {msg}")


def params(W, H, tau_inv, cs=0.0, model=MRT_LES, bc=CHANNEL, clamp=1, s_e=1.1, s_eps=1.1, s_q=1.1, sums=0,
           row_base=None, fast=0):
    """row_base: None, or the device address of a per-row base table (lbm_row_base) for 16-bit storage; fast: 1 for
    the hardware's approximate reciprocal and square root in the collision (lbm_params.fast)"""
    return LbmParams(int(W), int(H), int(model), int(bc), int(clamp), float(tau_inv), float(cs), float(s_e),
                     float(s_eps), float(s_q), int(sums), int(fast), row_base)


def row_base_table(ux, uy=None):
    """Host table of lbm_row_base (numpy float32 [H, ROW_BASE_STRIDE]) for a velocity per row, computed by the
    library's own host routine."""
    import numpy as np
    ux = np.ascontiguousarray(ux, np.float32)
    uy = None if uy is None else np.ascontiguousarray(uy, np.float32)
    out = np.zeros((ux.shape[0], ROW_BASE_STRIDE), np.float32)
    check(load().lbm_row_base(ux.shape[0], ux.ctypes.data_as(_vp), None if uy is None else uy.ctypes.data_as(_vp),
                              out.ctypes.data_as(_vp)), "lbm_row_base")
    return out
