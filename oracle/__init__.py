"""D2Q9 oracle -- test infrastructure only.

ctypes front end of ``oracle/liblbm_oracle.so`` (the C restatement in ``lbm_oracle.c`` /
``oracle_impl.h``) plus the independent numpy restatement in ``oracle/pyref.py``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this package, and only as the
checker: the product (``lbm-cuda_amd/``) never imports it.

Every function takes numpy arrays in the reference layout: populations ``f[q, y, x]`` (SoA, x fastest,
``lbm_smemaa_main.py:35,43``), ``mask[y, x]`` as uint8 (``lbm_smemaa_main.py:68``), ``inlet_u[y]``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.int64)
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.int64)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int64)
W_EXACT = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float64)

MRT, BGK = 0, 1
CHANNEL, PERIODIC = 0, 1


class OracleParams(ctypes.Structure):
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
    ]


def params(W, H, tau_inv, cs=0.0, model=MRT, bc=CHANNEL, clamp=1, s_e=1.1, s_eps=1.1, s_q=1.1, sums=0):
    return OracleParams(int(W), int(H), int(model), int(bc), int(clamp), float(tau_inv), float(cs),
                        float(s_e), float(s_eps), float(s_q), int(sums))


# -mfma: fma()/fmaf() become the hardware fused multiply-add instead of libm's software emulation (both are
# correctly rounded, so the results are the same bits); -ffp-contract=off still forbids contracting anything
# that is not written as an fma; -fno-math-errno drops the errno branch around sqrt (inputs are positive).
CFLAGS = ["-O2", "-mfma", "-fno-math-errno", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-Wall"]


def _host_has_fma():
    try:
        with open("/proc/cpuinfo") as fh:
            return any(line.startswith("flags") and " fma " in line + " " for line in fh)
    except OSError:
        return True  # no /proc: let the loader fail loudly if the instruction is missing


def build(verbose=False):
    """Compile the C restatement with gcc (no GPU, no reference sources involved)."""
    cmd = ["gcc"] + CFLAGS + ["-o", LIB_PATH, os.path.join(_HERE, "lbm_oracle.c"), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True, cwd=_HERE)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        if not _host_has_fma():
            raise RuntimeError("the oracle is built with -mfma and this host CPU has no FMA3 instructions")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(OracleParams)
        vp = ctypes.c_void_p
        for suf in ("_f64", "_f32"):
            getattr(L, "oracle_aa_step" + suf).argtypes = [P, vp, vp, vp, ctypes.c_int, vp, vp, vp]
            getattr(L, "oracle_pull_step" + suf).argtypes = [P, vp, vp, vp, vp, vp, vp, vp]
            getattr(L, "oracle_macro" + suf).argtypes = [P, vp, vp, vp, ctypes.c_int, vp, vp, vp]
            real = ctypes.c_double if suf == "_f64" else ctypes.c_float
            getattr(L, "oracle_collide" + suf).argtypes = [P, vp, ctypes.c_int, real, vp]
            getattr(L, "oracle_aa_step_rate" + suf).argtypes = [P, vp, vp, vp, ctypes.c_int, vp]
            getattr(L, "oracle_neq" + suf).argtypes = [P, vp, vp, vp, ctypes.c_int, vp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_neumaier_f32.argtypes = [vp, ctypes.c_int]
        L.oracle_neumaier_f32.restype = ctypes.c_float
        L.oracle_mrt_forward_f64.argtypes = [vp, vp]
        L.oracle_mrt_inverse_f64.argtypes = [vp, vp]
        L.oracle_diagnostics_f32.argtypes = [P, vp, vp, vp, ctypes.c_int, vp]
        L.oracle_aa_step_h16.argtypes = [P, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.oracle_rebase_window.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.oracle_convert_f32_h16.argtypes = [vp, vp, ctypes.c_long, ctypes.c_int]
        L.oracle_convert_h16_f32.argtypes = [vp, vp, ctypes.c_long, ctypes.c_int]
        _lib = L
    return _lib


def _suffix(dtype):
    if dtype == np.float64:
        return "_f64"
    if dtype == np.float32:
        return "_f32"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check_state(p, mask, f, inlet_u):
    W, H = p.W, p.H
    assert f.shape == (9, H, W) and f.flags.c_contiguous, f.shape
    assert mask.shape == (H, W) and mask.dtype == np.uint8 and mask.flags.c_contiguous
    assert inlet_u.shape == (H,) and inlet_u.dtype == f.dtype


def aa_step(p, mask, f, inlet_u, parity, macros=False):
    """One AA step in place on ``f`` (float32 = FP32 mirror, float64 = FP64 oracle)."""
    _check_state(p, mask, f, inlet_u)
    mac = [np.zeros((p.H, p.W), f.dtype) for _ in range(3)] if macros else [None] * 3
    getattr(lib(), "oracle_aa_step" + _suffix(f.dtype))(
        ctypes.byref(p), _ptr(mask), _ptr(f), _ptr(inlet_u), int(parity), _ptr(mac[1]), _ptr(mac[2]), _ptr(mac[0]))
    return tuple(mac) if macros else None


def aa_steps(p, mask, f, inlet_u, parity, n):
    """n AA steps in place; returns the parity after them."""
    for _ in range(n):
        aa_step(p, mask, f, inlet_u, parity)
        parity ^= 1
    return parity


def set_threads(n):
    """OpenMP threads of the oracle's loops (n <= 0: default); returns the count in effect."""
    return int(lib().oracle_set_threads(int(n)))


def aa_step_rate(p, mask, f, inlet_u, parity, w_cell):
    """One AA step in place whose MRT stress moments relax at w_cell[y, x] where that is >= 0 (the learned
    closure's rate); other cells follow the model's own rule.  w_cell has f's precision."""
    _check_state(p, mask, f, inlet_u)
    w_cell = np.ascontiguousarray(w_cell, dtype=f.dtype)
    getattr(lib(), "oracle_aa_step_rate" + _suffix(f.dtype))(
        ctypes.byref(p), _ptr(mask), _ptr(f), _ptr(inlet_u), int(parity), _ptr(w_cell))


def neq(p, mask, f, inlet_u, parity):
    """m7neq and m8neq of the pre-collision state of every updated cell, [2, H, W] (0 elsewhere)."""
    _check_state(p, mask, f, inlet_u)
    out = np.zeros((2, p.H, p.W), dtype=f.dtype)
    getattr(lib(), "oracle_neq" + _suffix(f.dtype))(ctypes.byref(p), _ptr(mask), _ptr(f), _ptr(inlet_u),
                                                     int(parity), _ptr(out))
    return out


def pull_step(p, mask, f_now, f_out, inlet_u):
    """Reference-compatible pull step (lbm_core4.cu:120-211); returns (rho, ux, uy)."""
    _check_state(p, mask, f_now, inlet_u)
    assert f_out.shape == f_now.shape and f_out.dtype == f_now.dtype
    mac = [np.zeros((p.H, p.W), f_now.dtype) for _ in range(3)]
    getattr(lib(), "oracle_pull_step" + _suffix(f_now.dtype))(
        ctypes.byref(p), _ptr(mask), _ptr(f_now), _ptr(f_out), _ptr(inlet_u), _ptr(mac[1]), _ptr(mac[2]),
        _ptr(mac[0]))
    return tuple(mac)


def macro(p, mask, f, inlet_u, parity):
    _check_state(p, mask, f, inlet_u)
    mac = [np.zeros((p.H, p.W), f.dtype) for _ in range(3)]
    getattr(lib(), "oracle_macro" + _suffix(f.dtype))(
        ctypes.byref(p), _ptr(mask), _ptr(f), _ptr(inlet_u), int(parity), _ptr(mac[1]), _ptr(mac[2]), _ptr(mac[0]))
    return tuple(mac)


def collide(p, f9, inlet=False, uin=0.0):
    """Collision of one cell; returns (post-collision f[9], (rho, ux, uy))."""
    g = np.array(f9, dtype=f9.dtype if isinstance(f9, np.ndarray) else np.float64).copy()
    mac = np.zeros(3, g.dtype)
    getattr(lib(), "oracle_collide" + _suffix(g.dtype))(ctypes.byref(p), _ptr(g), int(inlet), uin, _ptr(mac))
    return g, mac


def mrt_forward(f9):
    f = np.ascontiguousarray(f9, dtype=np.float64)
    m = np.zeros(9)
    lib().oracle_mrt_forward_f64(_ptr(f), _ptr(m))
    return m


def mrt_inverse(m9):
    m = np.ascontiguousarray(m9, dtype=np.float64)
    f = np.zeros(9)
    lib().oracle_mrt_inverse_f64(_ptr(m), _ptr(f))
    return f


def diagnostics(p, mask, f32, inlet_u32, parity):
    """FP64 (mass, jx, jy, Fx, Fy) of a float32 state."""
    _check_state(p, mask, f32, inlet_u32)
    out = np.zeros(5)
    lib().oracle_diagnostics_f32(ctypes.byref(p), _ptr(mask), _ptr(f32), _ptr(inlet_u32), int(parity), _ptr(out))
    return out


F16, BF16 = 1, 2
W_F32 = np.array([0.444444444] + [0.1111111111] * 4 + [0.0277777778] * 4, dtype=np.float32)


def to_h16(x32, fmt):
    """float32 -> IEEE half (fmt 1) or bfloat16 (fmt 2) bits, round to nearest even."""
    x = np.ascontiguousarray(x32, dtype=np.float32)
    y = np.empty(x.shape, np.uint16)
    lib().oracle_convert_f32_h16(_ptr(x), _ptr(y), x.size, int(fmt))
    return y


def from_h16(bits, fmt):
    b = np.ascontiguousarray(bits, dtype=np.uint16)
    y = np.empty(b.shape, np.float32)
    lib().oracle_convert_h16_f32(_ptr(b), _ptr(y), b.size, int(fmt))
    return y


def _slot_base(base, parity, shape, window=None):
    """FP32 base value of every slot [9, H, W] of a state at AA parity `parity` (slot p holds direction p, or opp(p)
    when reversed), from a base table [H, 16] and, inside a window (x0, y0, w, h, vel[h, w, 2]), the base values of
    the window's per-cell velocities; zeros without a table."""
    if base is None:
        return np.zeros(shape, np.float32)
    b = np.asarray(base, np.float32)
    dirs = OPP if parity else np.arange(9)
    out = np.ascontiguousarray(np.broadcast_to(b[:, dirs].T[:, :, None], shape))
    if window is not None:
        x0, y0, w, h, vel = window
        out[:, y0:y0 + h, x0:x0 + w] = np.moveaxis(vel_base(vel)[:, :, dirs], -1, 0)
    return out


def vel_base(vel):
    """Base values [..., 9] (float32) of window velocities vel[..., 2] (float16): w_q ((3 cu + 4.5 cu^2) - usq15),
    usq15 = 1.5 (ux^2 + uy^2), cu = c_q.u (one addition for the diagonals), in FP32 operations in the C oracle's
    order (vel_base_f32), so the values are the same bits."""
    v = np.asarray(vel, np.float16).astype(np.float32)
    ux, uy = v[..., 0], v[..., 1]
    f = np.float32
    usq15 = f(1.5) * (ux * ux + uy * uy)
    out = np.empty(v.shape[:-1] + (9,), np.float32)
    for q in range(9):
        if q == 0:
            cu = np.zeros_like(ux)
        elif CX[q] == 0:
            cu = uy if CY[q] > 0 else -uy
        elif CY[q] == 0:
            cu = ux if CX[q] > 0 else -ux
        else:
            cu = (ux if CX[q] > 0 else -ux) + (uy if CY[q] > 0 else -uy)
        out[..., q] = W_F32[q] * ((f(3.0) * cu + f(4.5) * (cu * cu)) - usq15)
    return out


def deviations(f64, fmt, base=None, parity=0, window=None):
    """16-bit state from an FP64 population array: h = f - w rounded to FP32 (w: the FP32 weight literals), minus
    the slot's base value (FP32) when a per-row base table [H, 16] (and a window) is given, rounded to 16 bits."""
    f64 = np.asarray(f64, np.float64)
    h = (f64 - W_F32.astype(np.float64)[:, None, None]).astype(np.float32)
    if base is not None:
        h = (h - _slot_base(base, parity, h.shape, window)).astype(np.float32)
    return to_h16(h, fmt)


def populations(h16, fmt, base=None, parity=0, window=None):
    """FP64 populations w + h of a 16-bit state (h = stored + the slot's base value, in FP32, with a base table)."""
    h = from_h16(h16, fmt)
    if base is not None:
        h = (h + _slot_base(base, parity, h.shape, window)).astype(np.float32)
    return h.astype(np.float64) + W_F32.astype(np.float64)[:, None, None]


class OracleWindow(ctypes.Structure):
    """Mirror of the C oracle's oracle_window: a window of per-cell base velocities (IEEE halves)."""

    _fields_ = [("x0", ctypes.c_int), ("y0", ctypes.c_int), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("vel", ctypes.c_void_p)]


def _window(window):
    """(x0, y0, w, h, vel[h, w, 2] float16) -> (OracleWindow, the velocity array kept alive), or (None, None)."""
    if window is None:
        return None, None
    x0, y0, w, h, vel = window
    vel = np.ascontiguousarray(vel, np.float16)
    assert vel.shape == (h, w, 2)
    return OracleWindow(int(x0), int(y0), int(w), int(h), vel.ctypes.data), vel


def rebase_window(W, H, h16, fmt, parity, window):
    """The kernels' window re-derivation on a 16-bit state in place; returns the new window velocities [h, w, 2]."""
    ow, keep = _window(window)
    nv = np.zeros(keep.shape, np.float16)
    lib().oracle_rebase_window(int(W), int(H), _ptr(h16), int(fmt), int(parity), ctypes.byref(ow), _ptr(nv))
    return nv


def aa_step_h16(p, mask, h16, fmt, inlet_u32, parity, macros=False, base=None, window=None):
    """One AA step in place on a 16-bit deviation state (FP32 collision, RNE storage); base: None or a per-row
    base table [H, 16] (float32) applied to every load and store; window: None or (x0, y0, w, h, vel[h, w, 2]),
    per-cell base velocities (float16) whose base values take precedence inside the window."""
    assert h16.dtype == np.uint16 and h16.shape == (9, p.H, p.W) and h16.flags.c_contiguous
    prof = np.ascontiguousarray(inlet_u32, dtype=np.float32)
    mac = [np.zeros((p.H, p.W), np.float32) for _ in range(3)] if macros else [None] * 3
    b = None
    if base is not None:
        b = np.ascontiguousarray(base, np.float32)
        assert b.shape == (p.H, 16)
    ow, keep = _window(window)
    lib().oracle_aa_step_h16(ctypes.byref(p), _ptr(mask), _ptr(h16), int(fmt), _ptr(prof), int(parity),
                             _ptr(mac[1]), _ptr(mac[2]), _ptr(mac[0]), None if b is None else _ptr(b),
                             None if ow is None else ctypes.byref(ow))
    return tuple(mac) if macros else None


# --------------------------------------------------------------------------------------------------
# State conventions shared by the tests
# --------------------------------------------------------------------------------------------------

def feq(rho, ux, uy):
    """Second-order equilibrium w rho (1 + 3cu + 4.5cu^2 - 1.5u^2) in FP64 (lbm_smemaa_main.py:35-40)."""
    rho = np.asarray(rho, np.float64)
    ux = np.asarray(ux, np.float64)
    uy = np.asarray(uy, np.float64)
    usq = ux * ux + uy * uy
    out = np.empty((9,) + np.broadcast(rho, ux, uy).shape)
    for q in range(9):
        cu = CX[q] * ux + CY[q] * uy
        out[q] = W_EXACT[q] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    return out


def observable(p, mask, parity, ghost=False):
    """Boolean [9, H, W] selecting the populations that carry state (stale slots excluded):
    updated cells only (no solids, wall rows or outlet ghost); in canonical form the inlet's q=1,5,8
    (which arrive from outside and are rebuilt by Zou-He) are excluded.  ghost=True also selects the
    outlet ghost column (rows 1..H-2) of a reversed state, which the next odd step reads."""
    H, W = p.H, p.W
    m = mask.astype(bool)
    upd = ~m
    if p.bc == CHANNEL:
        upd[0, :] = False
        upd[H - 1, :] = False
        upd[:, W - 1] = False
        upd[1:H - 1, 0] = True
    sel = np.broadcast_to(upd, (9, H, W)).copy()
    if p.bc == CHANNEL and parity == 0:
        for q in (1, 5, 8):
            sel[q, :, 0] = False
    if ghost and p.bc == CHANNEL and parity == 1:
        sel[:, 1:H - 1, W - 1] = True
    return sel


def reverse(f):
    """Read a reversed (odd-parity) state as post-collision populations: out[q] = f[opp q]."""
    return f[OPP]


def neumaier_f32(x):
    """Neumaier-compensated FP32 sum of x in index order: the summation of the collision with sums = 1."""
    import numpy as np
    a = np.ascontiguousarray(x, dtype=np.float32)
    return float(lib().oracle_neumaier_f32(a.ctypes.data_as(ctypes.c_void_p), int(a.size)))
