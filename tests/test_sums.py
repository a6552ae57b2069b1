"""CPU: the compensated (Neumaier) sums of the collision's sums = 1 mode.  The summation itself is pinned
against the exactly rounded sum (math.fsum) and a case plain FP32 summation gets wrong; the FP32 mirror in
that mode stays within the per-population FP64 gate."""
import math

import numpy as np
import pytest

import oracle as O
from lbm import geometry as G


def _round_f32(x):
    return float(np.float32(x))


def test_neumaier_matches_exact_sum():
    rng = np.random.default_rng(7)
    for n in (2, 3, 9, 17, 100):
        for _ in range(200):
            x = (rng.standard_normal(n) * 10.0 ** rng.integers(-8, 3, n)).astype(np.float32)
            exact = math.fsum(float(v) for v in x)
            got = O.neumaier_f32(x)
            ulp = abs(float(np.spacing(np.float32(exact)))) if exact != 0 else 1e-45
            assert abs(got - exact) <= 2 * ulp, (n, got, exact)


def test_neumaier_recovers_cancelled_terms():
    x = np.array([1e8, 1.0, -1e8, 1e-3], dtype=np.float32)
    naive = np.float32(0)
    for v in x:
        naive = np.float32(naive + v)
    assert float(naive) == pytest.approx(1e-3, abs=1e-9)  # the 1.0 is lost in plain FP32
    assert O.neumaier_f32(x) == _round_f32(math.fsum(float(v) for v in x))  # 1.001, correctly rounded


def test_mirror_with_compensated_sums_within_fp64_gate():
    W, H = 48, 24
    mask = G.channel_mask(W, H, cylinder=(12, 12, 3))
    prof = G.parabolic_profile(H, 0.05)
    f0 = G.channel_initial_state(W, H, prof, noise=1e-3, seed=3)
    for sums in (0, 1):
        p = O.params(W, H, 1.8, cs=0.12, sums=sums)
        f32 = np.array(f0, dtype=np.float32)
        f64 = np.array(f0, dtype=np.float64)
        par = 0
        for _ in range(6):
            prev64 = np.array(f32, dtype=np.float64)
            O.aa_step(p, mask, f32, prof.astype(np.float32), par)
            O.aa_step(O.params(W, H, 1.8, cs=0.12), mask, prev64, prof, par)  # FP64 from the same state
            sel = O.observable(p, mask, par ^ 1)
            rel = np.abs(f32[sel] - prev64[sel]) / np.maximum(np.abs(prev64[sel]), 1e-3)
            assert rel.max() < 1e-6, (sums, rel.max())
            par ^= 1


def _periodic_state(W, H, u_rms, drho, seed):
    rng = np.random.default_rng(seed)
    ux, uy = G.random_phase_velocity(W, H, 4, u_rms, seed=seed)
    rho = 1.0 + drho * rng.standard_normal((H, W))
    return np.array(G.feq(rho, ux, uy), dtype=np.float32)


def test_plain_sums_already_exact_at_low_mach():
    """In deviation form h_q = f_q - w_q is exact (Sterbenz) and lies on the grid of ulp(f_q) >= 2^-29; while the
    partial sums stay below 2^-5 they fit 24 bits, so the plain FP32 moment sums are exact and compensation
    changes nothing (u ~ 0.04, |rho - 1| ~ 1e-3)."""
    W, H = 32, 16
    mask = G.periodic_mask(W, H)
    a = _periodic_state(W, H, 0.04, 1e-3, 2)
    b = a.copy()
    z = np.zeros(H, np.float32)
    for par in (0, 1, 0, 1):
        O.aa_step(O.params(W, H, 1.7, cs=0.1, bc=O.PERIODIC), mask, a, z, par)
        O.aa_step(O.params(W, H, 1.7, cs=0.1, bc=O.PERIODIC, sums=1), mask, b, z, par)
    assert np.array_equal(a, b)


def test_compensated_sums_differ_when_sums_round():
    """Strong deviations (u ~ 0.25, |rho - 1| ~ 0.08): partial sums exceed 2^-5, plain sums round and the
    compensated mode gives different (last-bit) results."""
    W, H = 32, 16
    mask = G.periodic_mask(W, H)
    a = _periodic_state(W, H, 0.25, 0.08, 3)
    b = a.copy()
    z = np.zeros(H, np.float32)
    O.aa_step(O.params(W, H, 1.7, cs=0.1, bc=O.PERIODIC), mask, a, z, 0)
    O.aa_step(O.params(W, H, 1.7, cs=0.1, bc=O.PERIODIC, sums=1), mask, b, z, 0)
    assert not np.array_equal(a, b)
    assert np.abs(a - b).max() < 1e-6  # last-bit differences only
