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
    # the strip kernel has no base path (any k; without a configuration other k run as single steps)
    cfg = _lib.LbmFusedCfg(128, 8)
    for k in (2, 4):
        rc = L.lbm_step_fused_h16(ctypes.byref(p), fake, fake, ctypes.c_void_p(0x200000), O.F16, fake, 0, k,
                                  ctypes.byref(cfg), None, None, None, None)
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


# ------------------------------------------------------------------ per-cell base window
def _window_from_state(f64, parity, x0, y0, w, h):
    """Per-cell base velocities [h, w, 2] (float16) of a window: each cell's velocity in the FP64 state."""
    fd = np.asarray(f64, np.float64)[:, y0:y0 + h, x0:x0 + w]
    fd = fd[O.OPP] if parity else fd
    rho = fd.sum(0)
    return np.stack([(fd * O.CX[:, None, None]).sum(0) / rho, (fd * O.CY[:, None, None]).sum(0) / rho],
                    -1).astype(np.float16)


def _run16w(p, mask, f0, prof, steps, fmt, base, xywh, every):
    """16-bit mirror run with a base window (taken from the initial state) re-derived every `every` steps."""
    x0, y0, w, hh = xywh
    vel = _window_from_state(f0, 0, *xywh)
    h = O.deviations(f0, fmt, base=base, parity=0, window=(x0, y0, w, hh, vel))
    par = 0
    for n in range(steps):
        O.aa_step_h16(p, mask, h, fmt, prof.astype(np.float32), par, base=base, window=(x0, y0, w, hh, vel))
        par ^= 1
        if (n + 1) % every == 0:
            vel = O.rebase_window(p.W, p.H, h, fmt, par, (x0, y0, w, hh, vel))
    win = (x0, y0, w, hh, vel)
    return O.macro(p, mask, O.populations(h, fmt, base=base, parity=par, window=win), prof, par), win


def test_vel_base_values():
    """The window's base values from a velocity (the oracle's numpy restatement of the kernels' vel_base): the FP64
    equilibrium deviation to FP32 rounding error, with sum 0 and first moment u; opposite directions share cu
    up to sign, so a value and its opposite differ by exactly 2 w 3 cu in FP32 terms."""
    rng = np.random.default_rng(3)
    vel = rng.uniform(-0.12, 0.12, (7, 5, 2)).astype(np.float16)
    d = O.vel_base(vel)
    u = vel.astype(np.float64)
    for q in range(9):
        cu = O.CX[q] * u[..., 0] + O.CY[q] * u[..., 1]
        want = O.W_EXACT[q] * (3 * cu + 4.5 * cu * cu - 1.5 * (u ** 2).sum(-1))
        assert np.allclose(d[..., q], want, rtol=2e-6, atol=2e-9), q
    assert np.allclose(d.astype(np.float64).sum(-1), 0, atol=1e-8)
    assert np.allclose((d * O.CX).sum(-1), u[..., 0], atol=1e-7)
    assert np.allclose((d * O.CY).sum(-1), u[..., 1], atol=1e-7)
    assert not O.vel_base(np.zeros((2, 2), np.float16)).any()


def test_zero_velocity_window_equals_no_window():
    """A window of zero velocities over a rest-state row table: every base value is 0 either way, so the 16-bit
    mirror runs bit for bit as without the window."""
    W, H = 48, 24
    p = O.params(W, H, 1.6, cs=0.1)
    prof = G.parabolic_profile(H, 0.05)
    mask = G.channel_mask(W, H, cylinder=(12, 12, 3))
    base = _table(np.zeros(H))
    h1 = O.deviations(G.channel_initial_state(W, H, prof, 1e-3, 1), O.F16, base=base)
    h2 = h1.copy()
    win = (16, 6, 20, 12, np.zeros((12, 20, 2), np.float16))
    for par in (0, 1, 0, 1, 0, 1):
        O.aa_step_h16(p, mask, h1, O.F16, prof.astype(np.float32), par, base=base)
        O.aa_step_h16(p, mask, h2, O.F16, prof.astype(np.float32), par, base=base, window=win)
    assert np.array_equal(h1, h2)


def test_rebase_window_is_the_local_velocity():
    """lbm_rebase_window's re-derivation: the new per-cell velocity is the cell's decoded velocity rounded to half,
    the populations are kept to the 16-bit rounding of the new deviation, cells outside the window are untouched,
    and a second rebase barely moves the velocities."""
    W, H = 40, 30
    rng = np.random.default_rng(4)
    ux, uy = rng.uniform(-0.08, 0.08, (H, W)), rng.uniform(-0.08, 0.08, (H, W))
    f0 = O.feq(1.0 + rng.uniform(-0.01, 0.01, (H, W)), ux, uy)
    base = _table(np.full(H, 0.02))
    inside = np.zeros((H, W), bool)
    inside[7:18, 5:28] = True
    for fmt, half_ulp in ((O.F16, 2.0 ** -11), (O.BF16, 2.0 ** -8)):
        for parity in (0, 1):
            win = (5, 7, 23, 11, np.full((11, 23, 2), 0.02, np.float16))
            h = O.deviations(f0, fmt, base=base, parity=parity, window=win)
            before = O.populations(h, fmt, base=base, parity=parity, window=win)
            h_out = h.copy()
            nv = O.rebase_window(W, H, h, fmt, parity, win)
            after = O.populations(h, fmt, base=base, parity=parity, window=win[:4] + (nv,))
            assert np.array_equal(h[:, ~inside], h_out[:, ~inside])
            dev = np.abs(before - O.W_EXACT[:, None, None] - O._slot_base(base, parity, h.shape, win[:4] + (nv,)))
            assert (np.abs(after - before)[:, inside] <= half_ulp * dev[:, inside] + 1e-7).all()
            fd = before[O.OPP] if parity else before  # direction q sits in slot opp(q) when reversed
            rho = fd.sum(0)
            vx = ((fd * O.CX[:, None, None]).sum(0) / rho)[7:18, 5:28]
            vy = ((fd * O.CY[:, None, None]).sum(0) / rho)[7:18, 5:28]
            got = nv.astype(np.float64)
            assert np.allclose(got[..., 0], vx, rtol=2 ** -11, atol=1e-7), (fmt, parity)
            assert np.allclose(got[..., 1], vy, rtol=2 ** -11, atol=1e-7), (fmt, parity)
            nv2 = O.rebase_window(W, H, h.copy(), fmt, parity, win[:4] + (nv,))
            # the words' rounding moves the decoded velocity by ~half_ulp x 1e-2, plus the half's own rounding
            assert (np.abs(nv2.astype(np.float64) - got) <= 2.0 ** -9 * np.abs(got) + 16 * half_ulp * 1e-2).all()


def test_any_consistent_window_base_is_valid():
    """Any per-cell velocities give a valid window base as long as every load of a slot uses the base its store
    used (the slot's owner cell and direction): random velocities up to 0.1 keep the 16-bit run about as close to
    FP64 as the rest-state base, and re-derivations mid-run change nothing but rounding.  A load that took a
    neighbour's or another direction's base would be off by ~1e-2 per population and step; the same happens, and is
    detected, when the state is encoded against one table and stepped with another."""
    W, H, steps = 40, 32, 60
    p = O.params(W, H, 1.5, cs=0.1, bc=O.PERIODIC)
    mask = G.periodic_mask(W, H)
    mask[10:14, 12:15] = 1
    ux, uy = G.random_phase_velocity(W, H, 3, 0.04, seed=8)
    f0 = O.feq(np.ones((H, W)), ux, uy)
    prof = np.zeros(H)
    ref = _run64(p, mask, f0, prof, steps)
    sel = mask == 0
    base = _table(np.zeros(H))
    e_rest = _l2u(_run16(p, mask, f0, prof, steps, O.F16, base)[0], ref, sel)
    rng = np.random.default_rng(9)
    x0, y0, w, hh = 3, 5, 30, 20
    rv = rng.uniform(-0.1, 0.1, (hh, w, 2)).astype(np.float16)
    win = (x0, y0, w, hh, rv)

    def run(load_win, store_win, every=0):
        h = O.deviations(f0, O.F16, base=base, parity=0, window=store_win)
        cur = load_win
        for n in range(steps):
            O.aa_step_h16(p, mask, h, O.F16, prof.astype(np.float32), n & 1, base=base, window=cur)
            if every and (n + 1) % every == 0:
                cur = cur[:4] + (O.rebase_window(W, H, h, O.F16, (n + 1) & 1, cur),)
        return _l2u(O.macro(p, mask, O.populations(h, O.F16, base=base, parity=steps & 1, window=cur), prof,
                            steps & 1), ref, sel)

    e_rand = run(win, win)
    e_rebased = run(win, win, every=20)
    e_wrong = run(win[:4] + (rv[:, :, ::-1].copy(),), win)
    assert e_rand < 3 * e_rest and e_rebased < 3 * e_rest, (e_rand, e_rebased, e_rest)
    assert e_wrong > 30 * e_rest, (e_wrong, e_rest)


def test_cylinder_wake_window():
    """The channel flow past a cylinder with the inlet profile as the rows' base and a per-cell window over the
    near wake, re-derived every 10 steps: the velocity error against FP64 falls below three quarters of the row
    base's (1.1e-4 -> 6.1e-5 at this size and time; the larger share of the error at this Reynolds number and
    resolution is in the non-equilibrium part, which no base removes)."""
    W, H, steps = 256, 64, 400
    prof = G.parabolic_profile(H, 0.05)
    mask = G.channel_mask(W, H, cylinder=(64, 32, 8))
    f0 = G.channel_initial_state(W, H, prof, 1e-4, 2)
    p = O.params(W, H, float(np.float32(G.reynolds_tau_inv(0.05, 16, 200.0))), cs=0.1)
    ref = _run64(p, mask, f0, prof, steps)
    sel = O.observable(p, mask, 0)[0]
    base = _table(prof)
    e_rb = _l2u(_run16(p, mask, f0, prof, steps, O.F16, base)[0], ref, sel)
    m_w, _ = _run16w(p, mask, f0, prof, steps, O.F16, base, (48, 12, 144, 40), 10)
    e_w = _l2u(m_w, ref, sel)
    assert e_w < 0.75 * e_rb, (e_w, e_rb)


def test_window_argument_errors():
    L = lbm.load()
    base = np.zeros((8, 16), np.float32)
    wb = np.zeros((4, 4, 2), np.float16)
    fake = ctypes.c_void_p(0x100000)

    def rc_of(window, row_base=True):
        p = _lib.params(16, 8, 1.5, row_base=base.ctypes.data if row_base else None)
        p.window = ctypes.pointer(window)
        return L.lbm_step_aa_h16(ctypes.byref(p), fake, fake, O.F16, fake, 0, None, None, None, None)

    assert rc_of(_lib.LbmBaseWindow(2, 2, 4, 4, wb.ctypes.data), row_base=False) == -1
    assert b"row_base" in L.lbm_last_error()
    for bad in ((0, 2, 4, 4), (2, 2, 13, 4), (2, 5, 4, 4), (2, 2, 0, 4)):
        assert rc_of(_lib.LbmBaseWindow(*bad, wb.ctypes.data)) == -1, bad
        assert b"window" in L.lbm_last_error()
    assert rc_of(_lib.LbmBaseWindow(2, 2, 4, 4, wb.ctypes.data + 2)) == -1
    assert b"aligned" in L.lbm_last_error()
    p = _lib.params(16, 8, 1.5, row_base=base.ctypes.data)
    assert L.lbm_rebase_window(ctypes.byref(p), fake, O.F16, 0, fake, None) == -1
    assert b"window" in L.lbm_last_error()


# ------------------------------------------------------------------ the two-step kernel's window-first tile order
def _window_tile_rect(W, H, CH, window, CW=128):
    """launch_pair_fmt's rectangle of tiles whose region [x0 - 1, x0 + CW] x [y0 - 1, y0 + CH] meets the window."""
    nbx, nby = -(-W // CW), -(-H // CH)
    wx0, wy0, ww, wh = window
    bx0 = 0 if wx0 <= CW else (wx0 - 1) // CW
    by0 = 0 if wy0 <= CH else (wy0 - 1) // CH
    bx1, by1 = min(nbx - 1, (wx0 + ww) // CW), min(nby - 1, (wy0 + wh) // CH)
    return bx0, by0, bx1 - bx0 + 1, by1 - by0 + 1


def _pair_tile_order(b, nbx, nby, rect):
    """pair_tile_order (lbm_pair_tile.h): block b -> tile (bx, by), the rectangle's tiles first."""
    tx0, ty0, tw, th = rect
    nb, nw = nbx * nby, tw * th
    if b < nw:
        return tx0 + b % tw, ty0 + b // tw
    j, per = b - nw, (nb - nw) // 8
    t = (j % 8) * per + j // 8 if j < per * 8 else j
    above, rest = ty0 * nbx, nbx - tw
    if t >= above:
        k = t - above
        if k < th * rest:
            i = k % rest
            return (i if i < tx0 else i + tw), ty0 + k // rest
        t += nw
    return t % nbx, t // nbx


@pytest.mark.parametrize("W,H,CH", [(4096, 1024, 12), (300, 72, 8), (301, 90, 24), (131, 37, 16), (1024, 256, 12)])
def test_window_first_tile_order_is_a_permutation(W, H, CH):
    """The window's tiles (exactly those whose region meets it) come first, and the order is a permutation of all
    tiles, for windows anywhere in the lattice (edges, corners, the whole interior) and without one."""
    rng = np.random.default_rng(W + H + CH)
    nbx, nby = -(-W // 128), -(-H // CH)
    wins = [(1, 1, W - 3, H - 2), (1, 1, 1, 1), (W - 3, H - 2, 1, 1), (1, H // 2, W - 3, 1), (W // 2, 1, 1, H - 2)]
    for _ in range(20):
        x0, y0 = int(rng.integers(1, W - 2)), int(rng.integers(1, H - 1))
        wins.append((x0, y0, int(rng.integers(1, W - 1 - x0)), int(rng.integers(1, H - y0))))
    for win in wins:
        rect = _window_tile_rect(W, H, CH, win)
        meets = {(bx, by) for bx in range(nbx) for by in range(nby)
                 if bx * 128 + 128 >= win[0] and bx * 128 - 1 < win[0] + win[2]
                 and by * CH + CH >= win[1] and by * CH - 1 < win[1] + win[3]}
        order = [_pair_tile_order(b, nbx, nby, rect) for b in range(nbx * nby)]
        assert sorted(order) == [(bx, by) for bx in range(nbx) for by in range(nby)], win
        assert set(order[:rect[2] * rect[3]]) == meets, win
    # no window: the banded XCD order of xcd_tile
    per = nbx * nby // 8
    for b in range(nbx * nby):
        t = (b % 8) * per + b // 8 if b < per * 8 else b
        assert _pair_tile_order(b, nbx, nby, (0, 0, 0, 0)) == (t % nbx, t // nbx)


def test_obstacle_wake_window():
    """The wake window found from a channel mask's bodies equals the cylinder cases' own (4096 x 1024: the tile-
    aligned [1025, 1279) x [385, 635)); a channel without a body has none; a block's window starts at its centre."""
    from lbm import configs
    for case in (configs.cylinder_large(), configs.cylinder_scaled(1024, 256), configs.cylinder_small()):
        assert G.obstacle_wake_window(case.mask) == configs.wake_window(case), case.name
    assert configs.wake_window(configs.cylinder_large()) == (1025, 385, 254, 250)
    assert G.obstacle_wake_window(G.channel_mask(300, 64)) is None
    m = G.channel_mask(300, 64)
    m[20:31, 50:61] = 1  # an 11 x 11 block: centre (55, 25), half-height 5
    assert G.obstacle_wake_window(m) == (55, 15, 20, 20)
