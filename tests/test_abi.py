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
    # int32 W, H, model, bc, clamp; float tau_inv, cs, s_e, s_eps, s_q; int32 sums
    assert ctypes.sizeof(_lib.LbmParams) == 44
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
