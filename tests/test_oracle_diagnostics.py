"""CPU pins of the oracle's diagnostics and non-equilibrium moments, against what the mathematics fixes.

* ``oracle_diagnostics_f32`` (FP64 mass, momentum and momentum-exchange force of an FP32 state): on a periodic
  lattice with obstacles the force must balance the fluid's momentum change.  An even AA step collides in
  place (no streaming), so momentum is unchanged; an odd step streams twice -- the reversed state's
  post-collision populations into their neighbours, then the new post-collision populations into the next
  canonical state -- so J(after) - J(before) = -(F(before) + F(after)), where F of a reversed state is the force
  of the streaming that follows it and F of a canonical state the force of the streaming that produced it
  (lbm_core4.cu:129-138 half-way bounce-back: a population f*_q that meets a solid returns as f_opp(q), handing
  the wall 2 c_q f*_q).  A sign error, a missing factor 2 or the wrong slot breaks the balance by O(|F|).
* ``oracle_neq`` (m7 - m7eq, m8 - m8eq of the pre-collision state, lbm_core4.cu:172-173) against the
  independent numpy transcription ``pyref.neq`` for canonical and reversed states.
"""
import numpy as np
import pytest

import oracle as O
from oracle import pyref
from lbm import geometry as G


def _periodic_case(W=96, H=64, ux=0.05, uy=0.02, seed=7, symmetric=False):
    mask = G.periodic_mask(W, H)
    if symmetric:
        # a disc and a block, both symmetric about the row line y = (H - 1) / 2
        yy, xx = np.mgrid[0:H, 0:W]
        mask[((xx - 30) ** 2 + (yy - (H - 1) / 2) ** 2) <= 8 ** 2] = 1
        mask[H // 2 - 4:H - H // 2 + 4, 60:66] = 1
    else:
        yy, xx = np.mgrid[0:H, 0:W]
        mask[((xx - 30) ** 2 + (yy - 25) ** 2) <= 7 ** 2] = 1
        mask[40:47, 60:71] = 1
        mask[10:14, W - 3:W] = 1  # wraps across the periodic x edge
        mask[10:14, 0:2] = 1
        mask[H - 2:H, 80:84] = 1  # wraps across the periodic y edge
        mask[0:1, 80:84] = 1
    rng = np.random.default_rng(seed)
    u = np.full((H, W), ux) + (0 if symmetric else rng.uniform(-0.005, 0.005, (H, W)))
    v = np.full((H, W), uy) + (0 if symmetric else rng.uniform(-0.005, 0.005, (H, W)))
    f = O.feq(np.ones((H, W)), u, v)
    f[:, mask.astype(bool)] = O.feq(1.0, 0.0, 0.0)[:, None]
    return mask, np.ascontiguousarray(f)


def _op(W, H, cs=0.1):
    return O.params(W, H, 1.7, cs=cs, model=O.MRT, bc=O.PERIODIC)


def _diag(op, mask, f64, parity):
    return O.diagnostics(op, mask, np.ascontiguousarray(f64, dtype=np.float32), np.zeros(op.H, np.float32), parity)


@pytest.mark.parametrize("cs", [0.0, 0.1])
def test_momentum_exchange_balances_momentum_change(oracle_lib, cs):
    mask, f = _periodic_case()
    W, H = f.shape[2], f.shape[1]
    op = _op(W, H, cs)
    zero = np.zeros(H)
    seen = 0.0
    par = 0
    for step in range(12):
        before = _diag(op, mask, f, par)
        O.aa_step(op, mask, f, zero, par)
        par ^= 1
        after = _diag(op, mask, f, par)
        assert abs(after[0] - before[0]) < 1e-4, f"step {step}: mass {before[0]} -> {after[0]}"
        dj = after[1:3] - before[1:3]
        if par == 1:  # the even step just done collides in place: momentum unchanged, nothing streamed
            assert np.all(np.abs(dj) < 2e-5), f"step {step}: even step changed momentum by {dj}"
        else:
            force = before[3:5] + after[3:5]
            seen = max(seen, float(np.abs(force).max()))
            assert np.all(np.abs(dj + force) < 2e-5), f"step {step}: dJ {dj} vs -(F0 + F1) {-force}"
    assert seen > 0.05  # the balance is tested on a force far above the FP32 cast's error


def test_force_direction_and_symmetry(oracle_lib):
    """A flow along +x past obstacles symmetric about the mid row: drag along +x, no lift (to rounding), in
    both parities; the same flow along -x reverses the drag."""
    for sign in (1.0, -1.0):
        mask, f = _periodic_case(ux=0.05 * sign, uy=0.0, symmetric=True)
        W, H = f.shape[2], f.shape[1]
        op = _op(W, H)
        par = 0
        for _ in range(4):
            O.aa_step(op, mask, f, np.zeros(H), par)
            par ^= 1
            d = _diag(op, mask, f, par)
            assert d[3] * sign > 0.01, f"drag {d[3]} for flow sign {sign}"
            assert abs(d[4]) < 1e-6, f"lift {d[4]} on a symmetric case"


@pytest.mark.parametrize("parity", [0, 1])
@pytest.mark.parametrize("clamp", [True, False])
def test_neq_matches_independent_transcription(oracle_lib, parity, clamp):
    mask, f = _periodic_case(ux=0.03, uy=-0.02, seed=3)
    W, H = f.shape[2], f.shape[1]
    op = O.params(W, H, 1.7, cs=0.1, model=O.MRT, bc=O.PERIODIC, clamp=int(clamp))
    zero = np.zeros(H)
    for p in range(parity + 3):  # develop non-equilibrium parts first; end at the requested parity
        O.aa_step(op, mask, f, zero, p & 1)
    par = (parity + 3) & 1
    assert par == parity ^ 1
    O.aa_step(op, mask, f, zero, par)
    par ^= 1
    assert par == parity
    got = O.neq(op, mask, f, zero, par)
    want = pyref.neq(f, mask, par, clamp=clamp, periodic=True)
    scale = np.abs(want).max()
    assert scale > 1e-4
    assert np.allclose(got, want, rtol=0, atol=1e-12 * max(1.0, scale)), np.abs(got - want).max()
    # and the FP32 mirror's neq is the FP64 one to FP32 rounding
    got32 = O.neq(op, mask, f.astype(np.float32), zero.astype(np.float32), par)
    assert np.allclose(got32, want, rtol=0, atol=1e-6)


def test_neq_clamp_engaged(oracle_lib):
    """A state whose speeds exceed the clamp: the oracle's neq uses the clamped density and velocity, as the
    reference's Smagorinsky input does (lbm_core4.cu:142-144,172-173)."""
    W, H = 40, 24
    rng = np.random.default_rng(11)
    mask = G.periodic_mask(W, H)
    f = np.ascontiguousarray(O.feq(rng.uniform(0.3, 1.8, (H, W)), rng.uniform(-0.6, 0.6, (H, W)),
                                   rng.uniform(-0.6, 0.6, (H, W))))
    op = O.params(W, H, 1.7, cs=0.1, model=O.MRT, bc=O.PERIODIC, clamp=1)
    got = O.neq(op, mask, f, np.zeros(H), 0)
    want = pyref.neq(f, mask, 0, clamp=True, periodic=True)
    unclamped = pyref.neq(f, mask, 0, clamp=False, periodic=True)
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    assert np.abs(want - unclamped).max() > 1e-3  # the clamps are engaged on this state
