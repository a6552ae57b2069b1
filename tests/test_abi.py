"""CPU checks of the C-ABI boundary: liblbm.so loads, exports every function include/lbm.h declares,
the ctypes structure matches the C layout, and argument errors come back as codes (no GPU call made)."""
import ctypes
import os
import re

import pytest

import lbm
from lbm import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("lbm_step_aa", "lbm_step_pull", "lbm_step_fused", "lbm_macro", "lbm_visualize",
                 "lbm_diagnostics", "lbm_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = lbm.load()
    for name in declared_functions():
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_params_struct_layout():
    # int32 W, H, model, bc, clamp; float tau_inv, cs, s_e, s_eps, s_q; int32 sums, fast; const float *row_base
    assert ctypes.sizeof(_lib.LbmParams) == 64
    assert _lib.LbmParams.fast.offset == 44 and _lib.LbmParams.row_base.offset == 48
    assert _lib.LbmParams.window.offset == 56 and ctypes.sizeof(_lib.LbmBaseWindow) == 24
    assert _lib.LbmParams.tau_inv.offset == 20
    assert _lib.LbmParams.sums.offset == 40
    import oracle as O
    assert ctypes.sizeof(O.OracleParams) == 44 and O.OracleParams.sums.offset == 40


def test_sums_argument_checked():
    L = lbm.load()
    p = _lib.params(16, 8, 1.5, sums=2)
    assert L.lbm_step_aa(ctypes.byref(p), None, None, None, 0, None, None, None, None) == -1
    assert b"sums" in L.lbm_last_error()


def test_argument_errors_return_codes():
    L = lbm.load()
    bad = _lib.params(2, 2, 1.5)
    rc = L.lbm_step_aa(ctypes.byref(bad), None, None, None, 0, None, None, None, None)
    assert rc == -2
    assert b"too small" in L.lbm_last_error()
    big = _lib.params(40000, 40000, 1.5)
    assert L.lbm_step_aa(ctypes.byref(big), None, None, None, 0, None, None, None, None) == -2
    p = _lib.params(16, 8, 2.5)
    assert L.lbm_step_aa(ctypes.byref(p), None, None, None, 0, None, None, None, None) == -1
    p = _lib.params(16, 8, 1.5)
    assert L.lbm_step_aa(ctypes.byref(p), None, None, None, 0, None, None, None, None) == -1  # NULL mask
    with pytest.raises(lbm.LbmError):
        _lib.check(L.lbm_step_aa(ctypes.byref(p), None, None, None, 0, None, None, None, None), "lbm_step_aa")


def test_fused_max_k():
    assert lbm.load().lbm_fused_max_k() == 8


# Kernels addressing populations through buffer resources keep the record size 9*W*H*bytes and every byte
# offset in unsigned 32 bits, so FP32 lattices stop at W*H < 2^32 / 36 = 119 304 647.1 cells and 16-bit
# ones at the element limit 9*W*H < 2^31 (238 609 294 cells).
FP32_MAX_CELLS = (2 ** 32 - 1) // 36          # 119 304 647
H16_MAX_CELLS = (2 ** 31 - 1) // 9            # 238 609 294


def _size_rc(L, fn, W, H, *rest):
    p = _lib.params(W, H, 1.5)
    return getattr(L, fn)(ctypes.byref(p), *rest)


def test_fp32_byte_offset_limit():
    """16384 x 8192 FP32 (2^27 cells, 4.8 GB) overflows 32-bit byte offsets: every FP32 entry point that
    addresses f through buffer resources refuses it with LBM_ERR_SIZE before any argument check."""
    L = lbm.load()
    W, H = 16384, 8192
    assert 9 * W * H * 4 >= 2 ** 32 and 9 * W * H < 2 ** 31
    assert _size_rc(L, "lbm_step_aa", W, H, None, None, None, 0, None, None, None, None) == -2
    assert b"2^32" in L.lbm_last_error()
    assert _size_rc(L, "lbm_step_aa_n", W, H, None, None, None, 0, 1, None) == -2
    assert _size_rc(L, "lbm_step_fused", W, H, None, None, None, None, 0, 2, None, None, None, None) == -2
    assert _size_rc(L, "lbm_step_fused_cfg", W, H, None, None, None, None, 0, 2, None, None, None, None,
                    None) == -2
    assert _size_rc(L, "lbm_step_fused_n", W, H, None, None, None, None, 0, 2, 1, None, None) == -2
    # the same lattice with 16-bit deviations passes the size check and fails on the NULL buffers instead
    assert _size_rc(L, "lbm_step_aa_h16", W, H, None, None, 1, None, 0, None, None, None, None) == -1
    assert _size_rc(L, "lbm_step_fused_h16", W, H, None, None, None, 1, None, 0, 2, None, None, None, None,
                    None) == -1
    # the closure steps FP32 populations through buffer resources too
    assert _size_rc(L, "lbm_neq", W, H, None, None, 0, None, None) == -2


def test_fp32_byte_offset_limit_boundary():
    """One row either side of the limit: the largest FP32 lattice accepted, the next one refused."""
    L = lbm.load()
    W = 16384
    H_ok = FP32_MAX_CELLS // W                    # 7281 rows: 9*W*H*4 < 2^32
    assert 9 * W * H_ok * 4 < 2 ** 32 <= 9 * W * (H_ok + 1) * 4
    assert _size_rc(L, "lbm_step_aa", W, H_ok, None, None, None, 0, None, None, None, None) == -1  # NULL mask
    assert _size_rc(L, "lbm_step_aa", W, H_ok + 1, None, None, None, 0, None, None, None, None) == -2
    # 16-bit: the element limit decides
    W2 = 32768
    H2 = H16_MAX_CELLS // W2
    assert _size_rc(L, "lbm_step_aa_h16", W2, H2, None, None, 1, None, 0, None, None, None, None) == -1
    assert _size_rc(L, "lbm_step_aa_h16", W2, H2 + 1, None, None, 1, None, 0, None, None, None, None) == -2


def test_ensemble_rejects_oversized_fp32_member():
    from lbm.ensemble import LbmInstance
    L = lbm.load()
    out = ctypes.c_void_p()
    inst = (LbmInstance * 1)()
    inst[0] = LbmInstance(_lib.params(16384, 8192, 1.5, bc=_lib.PERIODIC), 1, 1, 0)
    assert L.lbm_ensemble_create(inst, 1, _lib.FMT_F32, ctypes.byref(out)) == -2
    assert not out.value


def test_fused_default_cfg_validates():
    L = lbm.load()
    cfg = _lib.LbmFusedCfg()
    p = _lib.params(4096, 1024, 1.5)
    assert L.lbm_fused_default_cfg(ctypes.byref(p), 2, ctypes.byref(cfg)) == 0
    assert cfg.strip == 128 and cfg.chunk >= 1
    for k in (0, 9, 63, 100):
        assert L.lbm_fused_default_cfg(ctypes.byref(p), k, ctypes.byref(cfg)) == -1
    small = _lib.params(2, 2, 1.5)
    assert L.lbm_fused_default_cfg(ctypes.byref(small), 2, ctypes.byref(cfg)) == -2


def test_fused_h16_rejects_strip_64():
    """16-bit strip kernels exist for 128-column strips only: 64 is refused, not rewritten."""
    L = lbm.load()
    p = _lib.params(64, 32, 1.5)
    cfg = _lib.LbmFusedCfg(64, 16)
    # non-NULL dummy pointers that are never dereferenced: the configuration check comes first
    m, a, b, u = 0x1000, 0x100000, 0x200000, 0x300000
    rc = L.lbm_step_fused_h16(ctypes.byref(p), m, a, b, _lib.FMT_F16, u, 0, 2, ctypes.byref(cfg), None, None,
                              None, None)
    assert rc == -3 and b"strip" in L.lbm_last_error()


def test_ensemble_and_graph_argument_errors():
    """Ensemble / graph entry points reject bad arguments before touching the GPU."""
    from lbm.ensemble import LbmInstance
    L = lbm.load()
    out = ctypes.c_void_p()
    assert L.lbm_ensemble_create(None, 0, 0, ctypes.byref(out)) == -1
    inst = (LbmInstance * 2)()
    inst[0] = LbmInstance(_lib.params(64, 32, 1.5), 1, 1, 1)
    inst[1] = LbmInstance(_lib.params(64, 32, 1.5), 0, 1, 1)  # NULL mask
    assert L.lbm_ensemble_create(inst, 2, 0, ctypes.byref(out)) == -1
    assert b"instance 1" in L.lbm_last_error()
    inst[1] = LbmInstance(_lib.params(64, 32, 1.5, model=_lib.BGK), 1, 1, 1)
    assert L.lbm_ensemble_create(inst, 2, 0, ctypes.byref(out)) == -1
    assert b"differs" in L.lbm_last_error()
    assert L.lbm_ensemble_create(inst, 1, 7, ctypes.byref(out)) == -1  # bad format
    assert not out.value
    assert L.lbm_ensemble_step(None, 0, 0, None) == -1
    assert L.lbm_ensemble_info(None, None, None, None) == -1
    assert L.lbm_ensemble_destroy(None) == 0
    assert L.lbm_graph_begin(None) == -1  # the default stream cannot be captured
    assert L.lbm_graph_launch(None, None) == -1
    assert L.lbm_graph_end(None, None) == -1
    assert L.lbm_graph_destroy(None) == 0
