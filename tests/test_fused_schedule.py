"""CPU check of the temporally blocked kernel's schedule (tests/fused_emul.py replays k_fused's sweep with
bounds and row-disjointness assertions) against k FP64 oracle AA steps."""
import numpy as np
import pytest

import fused_emul as E
import oracle as O
from lbm import geometry as G

F32 = lambda v: float(np.float32(v))  # noqa: E731


def _case(W, H, periodic, par, op):
    if periodic:
        mask = G.periodic_mask(W, H)
        mask[H // 3:H // 3 + 3, W // 2:W // 2 + 4] = 1
        ux, uy = G.random_phase_velocity(W, H, 3, 0.04, seed=1)
        f0 = G.feq(np.ones((H, W)), ux, uy)
        prof = np.zeros(H)
    else:
        mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 3))
        prof = G.parabolic_profile(H, 0.05)
        f0 = G.channel_initial_state(W, H, prof, noise=1e-3, seed=2)
    f = f0.copy()
    if par == 1:
        O.aa_step(op, mask, f, prof, 0)
    return mask, prof, f


@pytest.mark.parametrize("K", [1, 2, 3, 8])
@pytest.mark.parametrize("par", [0, 1])
@pytest.mark.parametrize("periodic,model,strip,chunk", [(False, O.MRT, 128, None), (True, O.MRT, 128, None),
                                                        (False, O.BGK, 128, None), (False, O.MRT, 64, 24),
                                                        (True, O.MRT, 64, 7)])
def test_fused_schedule_matches_single_steps(oracle_lib, K, par, periodic, model, strip, chunk):
    W, H = (150, 40) if not periodic else (140, 36)
    cs = 0.1 if model == O.MRT else 0.0
    op = O.params(W, H, 1.7, cs=cs, model=model, bc=O.PERIODIC if periodic else O.CHANNEL)
    mask, prof, f = _case(W, H, periodic, par, op)
    ref = f.copy()
    rpar = O.aa_steps(op, mask, ref, prof, par, K)
    p = dict(tau_inv=F32(1.7), cs=F32(cs), model=model, clamp=1, s_e=F32(1.1), s_eps=F32(1.1), s_q=F32(1.1))
    out = E.fused(f.copy(), mask, prof, par, K, p, periodic, strip=strip, chunk=chunk)
    sel = O.observable(op, mask, rpar)
    assert not np.isnan(out[sel]).any()
    assert np.abs(out - ref)[sel].max() < 1e-13


@pytest.mark.parametrize("K,strip,chunk", [(3, 64, 24), (1, 128, 16), (2, 64, 9), (5, 128, 40)])
def test_fused_chained_launches_keep_the_ghost(oracle_lib, K, strip, chunk):
    """Consecutive launches (each reading the previous one's output) match single steps, including the
    outlet ghost of reversed states; W = 301 puts strip core boundaries between W-2 and W-1."""
    W, H = 301, 40
    op = O.params(W, H, 1.7, cs=0.1)
    mask, prof, f = _case(W, H, False, 0, op)
    p = dict(tau_inv=F32(1.7), cs=F32(0.1), model=O.MRT, clamp=1, s_e=F32(1.1), s_eps=F32(1.1), s_q=F32(1.1))
    ref = f.copy()
    par = 0
    for launch in range(3):
        out = E.fused(f, mask, prof, par, K, p, False, strip=strip, chunk=chunk)
        rpar = O.aa_steps(op, mask, ref, prof, par, K)
        sel = O.observable(op, mask, rpar, ghost=True)
        assert np.abs(out - ref)[sel].max() < 1e-13, f"launch {launch}"
        f = np.where(np.isnan(out), 0.0, out)
        par = rpar


@pytest.mark.parametrize("periodic,model,W,H,CH", [(False, O.MRT, 150, 40, 8), (True, O.MRT, 140, 36, 16),
                                                   (False, O.BGK, 131, 19, 8), (True, O.MRT, 40, 12, 8),
                                                   (False, O.MRT, 70, 9, 16)])
def test_pair_schedule_matches_two_single_steps(oracle_lib, periodic, model, W, H, CH):
    """The two-step tile kernel's phases (tests/fused_emul.py: pair) reproduce two FP64 oracle AA steps on
    every observable slot, chained over three launches, and no output slot is written twice -- the
    single-writer rule that lets the odd phase scatter straight to HBM."""
    cs = 0.1 if model == O.MRT else 0.0
    op = O.params(W, H, 1.7, cs=cs, model=model, bc=O.PERIODIC if periodic else O.CHANNEL)
    mask, prof, f = _case(W, H, periodic, 0, op)
    p = dict(tau_inv=F32(1.7), cs=F32(cs), model=model, clamp=1, s_e=F32(1.1), s_eps=F32(1.1), s_q=F32(1.1))
    ref = f.copy()
    for launch in range(3):
        out, written = E.pair(f, mask, prof, p, periodic, CH)
        assert written.max() <= 1, "an output slot has two writers"
        O.aa_steps(op, mask, ref, prof, 0, 2)
        sel = O.observable(op, mask, 0)
        assert not np.isnan(out[sel]).any(), f"launch {launch}: observable slot not written"
        assert np.abs(out - ref)[sel].max() < 1e-13, f"launch {launch}"
        f = np.where(np.isnan(out), 1e30, out)  # unwritten slots hold garbage: the next launch must not read them


@pytest.mark.parametrize("periodic,model,W,H,CH", [(False, O.MRT, 150, 40, 8), (True, O.MRT, 140, 36, 16),
                                                   (False, O.BGK, 131, 19, 8), (True, O.MRT, 40, 12, 8),
                                                   (False, O.MRT, 70, 9, 16), (False, O.MRT, 300, 50, 12),
                                                   (True, O.BGK, 136, 30, 12), (False, O.MRT, 150, 60, 24),
                                                   (True, O.MRT, 90, 50, 20), (False, O.BGK, 70, 40, 32),
                                                   (True, O.MRT, 40, 12, 24), (False, O.MRT, 150, 44, 14),
                                                   (True, O.MRT, 70, 30, 14)])
def test_quad_schedule_matches_four_single_steps(oracle_lib, periodic, model, W, H, CH):
    """The four-step tile kernel's phases (tests/fused_emul.py: quad) -- the even step of the 3-ring region,
    the odd step of ring 2 and the even step of ring 1 in place in LDS, the core's odd step scattered out --
    reproduce four FP64 oracle AA steps on every observable slot, chained over three launches; the in-place
    phases are race free and no output slot has two writers."""
    cs = 0.1 if model == O.MRT else 0.0
    op = O.params(W, H, 1.7, cs=cs, model=model, bc=O.PERIODIC if periodic else O.CHANNEL)
    mask, prof, f = _case(W, H, periodic, 0, op)
    p = dict(tau_inv=F32(1.7), cs=F32(cs), model=model, clamp=1, s_e=F32(1.1), s_eps=F32(1.1), s_q=F32(1.1))
    ref = f.copy()
    for launch in range(3):
        out, written = E.quad(f, mask, prof, p, periodic, CH)
        assert written.max() <= 1, "an output slot has two writers"
        O.aa_steps(op, mask, ref, prof, 0, 4)
        sel = O.observable(op, mask, 0)
        assert not np.isnan(out[sel]).any(), f"launch {launch}: observable slot not written"
        assert np.abs(out - ref)[sel].max() < 1e-13, f"launch {launch}"
        f = np.where(np.isnan(out), 1e30, out)  # unwritten slots hold garbage: the next launch must not read them
