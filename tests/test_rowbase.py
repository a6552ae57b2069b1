"""CPU: 16-bit storage relative to a per-row base state.

The library's host table routine (lbm_row_base) against the equilibrium formula, the argument errors of the entry
points that do not take a base (no GPU call is made), and the oracle's 16-bit mirror with a base table pinned
against the FP64 oracle on flows where the base is exact or close:
  - a uniform flow whose base is that flow: the stored words are the equilibrium's last-bit residue, so the
    16-bit run keeps the FP64 velocity to a few 1e-8 (the rest-state base keeps it only to ~1e-5);
  - a periodic (decaying) shear flow ux(y) with its initial profile as the base: the error is a fifth of the
    rest-state base's or less -- a load that took another row's or direction's value than its store (the gather's
    row y - c_qy, the scatter's y + c_qy) would add w_q 3 c_q.du ~ 1e-3 per population and step and fail this by
    orders of magnitude;
  - the channel flow past a cylinder with the inlet profile as the base: bounce-back slots change direction and
    take this row's value of the reflected direction, and a mistake there adds ~w_q 6 c_q.u at the body.
"""
import ctypes

import numpy as np
import pytest

import lbm
import oracle as O
from lbm import _lib
from lbm import geometry as G


def _table(ux, uy=None):
    return _lib.row_base_table(np.asarray(ux, np.float32), None if uy is None else np.asarray(uy, np.float32))


def test_row_base_table_formula():
    rng = np.random.default_rng(0)
    ux = rng.uniform(-0.1, 0.1, 37).astype(np.float32)
    uy = rng.uniform(-0.1, 0.1, 37).astype(np.float32)
    t = _table(ux, uy)
    assert t.shape == (37, 16) and t.dtype == np.float32
    u, v = ux.astype(np.float64), uy.astype(np.float64)
    for q in range(9):
        cu = O.CX[q] * u + O.CY[q] * v
        want = (O.W_EXACT[q] * (3 * cu + 4.5 * cu * cu - 1.5 * (u * u + v * v))).astype(np.float32)
        assert np.array_equal(t[:, q], want), q
    assert not t[:, 9:].any()
    # the table is the equilibrium's deviation from w: sum_q d_q = 0, sum_q c_q d_q = u (to FP32 rounding)
    assert np.allclose(t[:, :9].astype(np.float64).sum(1), 0.0, atol=1e-8)
    assert np.allclose((t[:, :9].astype(np.float64) * O.CX).sum(1), u, atol=1e-8)
    assert np.allclose((t[:, :9].astype(np.float64) * O.CY).sum(1), v, atol=1e-8)
    t0 = _table(ux)  # uy = NULL means 0
    assert np.array_equal(t0, _table(ux, np.zeros_like(ux)))


def test_row_base_argument_errors():
    L = lbm.load()
    assert L.lbm_row_base(0, None, None, None) == -1
    base = np.zeros((8, 16), np.float32)
    p = _lib.params(16, 8, 1.5, row_base=base.ctypes.data + 4)  # misaligned
    fake = ctypes.c_void_p(0x100000)
    rc = L.lbm_step_aa_h16(ctypes.byref(p), fake, fake, O.F16, fake, 0, None, None, None, None)
    assert rc == -1 and b"aligned" in L.lbm_last_error()
    p = _lib.params(16, 8, 1.5, row_base=base.ctypes.data)
    rc = L.lbm_convert(ctypes.byref(p), fake, _lib.FMT_F16, ctypes.c_void_p(0x200000), _lib.FMT_F32, None)
    assert rc == -1 and b"lbm_convert_state" in L.lbm_last_error()
    rc = L.lbm_convert_state(ctypes.byref(p), fake, _lib.FMT_F16, ctypes.c_void_p(0x200000), _lib.FMT_F32, 2, None)
    assert rc == -1 and b"parity" in L.lbm_last_error()
    # the strip kernel and k != 2 have no base path
    cfg = _lib.LbmFusedCfg(128, 8)
    rc = L.lbm_step_fused_h16(ctypes.byref(p), fake, fake, ctypes.c_void_p(0x200000), O.F16, fake, 0, 2,
                              ctypes.byref(cfg), None, None, None, None)
    assert rc == -3 and b"row-based" in L.lbm_last_error()
    rc = L.lbm_step_fused_h16(ctypes.byref(p), fake, fake, ctypes.c_void_p(0x200000), O.F16, fake, 0, 4, None, None,
                              None, None, None)
    assert rc == -3 and b"row-based" in L.lbm_last_error()
    # ensembles: rejected while the member table is checked, before anything is allocated
    from lbm.ensemble import LbmInstance
    inst = (LbmInstance * 1)()
    inst[0].p = p
    out = ctypes.c_void_p()
    rc = L.lbm_ensemble_create(inst, 1, _lib.FMT_F16, ctypes.byref(out))
    assert rc == -3 and b"row_base" in L.lbm_last_error() and not out.value


def _run16(p, mask, f0, prof, steps, fmt, base):
    h = O.deviations(f0, fmt, base=base, parity=0)
    par = 0
    for _ in range(steps):
        O.aa_step_h16(p, mask, h, fmt, prof.astype(np.float32), par, base=base)
        par ^= 1
    return O.macro(p, mask, O.populations(h, fmt, base=base, parity=par), prof, par), par


def _run64(p, mask, f0, prof, steps):
    f = f0.astype(np.float64).copy()
    par = O.aa_steps(p, mask, f, prof, 0, steps)
    return O.macro(p, mask, f, prof, par)


def _l2u(m, ref, sel):
    _, a, b = (np.asarray(v, np.float64) for v in m)
    _, A, B = (np.asarray(v, np.float64) for v in ref)
    return float(np.sqrt((((a - A) ** 2 + (b - B) ** 2)[sel]).sum() / ((A ** 2 + B ** 2)[sel]).sum()))


def test_uniform_flow_with_its_own_base_is_exact():
    W, H, steps = 24, 16, 40
    u, v = 0.05, -0.02
    p = O.params(W, H, 1.6, cs=0.1, bc=O.PERIODIC)
    mask = G.periodic_mask(W, H)
    f0 = O.feq(np.ones((H, W)), np.full((H, W), u), np.full((H, W), v))
    prof = np.zeros(H)
    ref = _run64(p, mask, f0, prof, steps)
    sel = mask == 0
    base = _table(np.full(H, u), np.full(H, v))
    e_rb = _l2u(_run16(p, mask, f0, prof, steps, O.F16, base)[0], ref, sel)
    e_w = _l2u(_run16(p, mask, f0, prof, steps, O.F16, None)[0], ref, sel)
    assert e_rb < 2e-6 and e_w > 1e-5, (e_rb, e_w)
    # a base for the opposite directions (a plausible slip) is no better than the rest state
    bad = base.copy()
    bad[:, :9] = base[:, O.OPP]
    assert _l2u(_run16(p, mask, f0, prof, steps, O.F16, bad)[0], ref, sel) > 50 * e_rb


@pytest.mark.parametrize("fmt", [O.F16, O.BF16])
def test_shear_flow_with_its_profile_as_base(fmt):
    W, H, steps = 32, 64, 300
    y = np.arange(H)
    ux = 0.05 * np.sin(2 * np.pi * (y + 0.5) / H)
    p = O.params(W, H, 1.2, bc=O.PERIODIC)
    mask = G.periodic_mask(W, H)
    rng = np.random.default_rng(1)
    f0 = O.feq(np.ones((H, W)), ux[:, None] + rng.uniform(-1e-4, 1e-4, (H, W)), rng.uniform(-1e-4, 1e-4, (H, W)))
    prof = np.zeros(H)
    ref = _run64(p, mask, f0, prof, steps)
    sel = mask == 0
    base = _table(ux)
    e_rb = _l2u(_run16(p, mask, f0, prof, steps, fmt, base)[0], ref, sel)
    e_w = _l2u(_run16(p, mask, f0, prof, steps, fmt, None)[0], ref, sel)
    assert e_rb < e_w / 5, (e_rb, e_w)
    # any table is a valid base as long as loads and stores agree on it: the profile shifted by a row is about as
    # good (only the stored magnitudes change), while a load that used another row's value than the store did would
    # be off by w_q 3 c_q.du ~ 1e-3 per population and step
    e_shift = _l2u(_run16(p, mask, f0, prof, steps, fmt, np.roll(base, 1, axis=0))[0], ref, sel)
    assert e_shift < e_w / 4, (e_shift, e_w)


def test_cylinder_channel_with_inlet_base():
    W, H, steps = 256, 64, 400
    prof = G.parabolic_profile(H, 0.05)
    mask = G.channel_mask(W, H, cylinder=(64, 32, 8))
    f0 = G.channel_initial_state(W, H, prof, 1e-4, 2)
    p = O.params(W, H, float(np.float32(G.reynolds_tau_inv(0.05, 16, 200.0))), cs=0.1)
    ref = _run64(p, mask, f0, prof, steps)
    sel = O.observable(p, mask, 0)[0]
    base = _table(prof)
    (m_rb, par) = _run16(p, mask, f0, prof, steps, O.F16, base)
    e_rb = _l2u(m_rb, ref, sel)
    e_w = _l2u(_run16(p, mask, f0, prof, steps, O.F16, None)[0], ref, sel)
    assert par == 0 and e_rb < e_w / 4, (e_rb, e_w)
