"""CPU checks of the 16-bit deviation storage in the oracle: conversions round to nearest even (IEEE half
against numpy, bfloat16 against its definition), and the FP16/BF16 runs track FP64 within the precision
the formats allow."""
import numpy as np
import pytest

import oracle as O
from lbm import geometry as G


def test_half_conversion_matches_numpy(oracle_lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1e-3, 200000), rng.normal(0, 0.3, 200000), rng.uniform(-1e-7, 1e-7, 50000),
                        [0.0, -0.0, 6.1035156e-05, 5.9604645e-08, 65504.0, 70000.0, -1e-9]]).astype(np.float32)
    bits = O.to_h16(x, O.F16)
    assert np.array_equal(bits, x.astype(np.float16).view(np.uint16))
    back = O.from_h16(bits, O.F16)
    assert np.array_equal(back, x.astype(np.float16).astype(np.float32))


def test_bf16_conversion_round_to_nearest_even(oracle_lib):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1e-2, 200000), rng.normal(0, 10, 1000)]).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    # exact RNE from the definition: compare the discarded 16 bits with half an ulp
    keep = u >> 16
    rem = u & 0xFFFF
    up = (rem > 0x8000) | ((rem == 0x8000) & ((keep & 1) == 1))
    want = (keep + up).astype(np.uint16)
    assert np.array_equal(O.to_h16(x, O.BF16), want)
    ties = np.array([0x3F808000, 0x3F818000], np.uint32).view(np.float32)  # ties to even: down, up
    assert list(O.to_h16(ties, O.BF16)) == [0x3F80, 0x3F82]


@pytest.mark.parametrize("fmt,tol", [(O.F16, 2e-3), (O.BF16, 2e-2)])
def test_h16_run_tracks_fp64(oracle_lib, fmt, tol):
    """Periodic shear flow, 200 steps: L2 error of u for FP16/BF16 deviation storage vs the FP64 oracle."""
    W, H = 96, 64
    ux, uy = G.shear_layer_velocity(W, H, delta=20.0, noise=0.0)
    f0 = G.feq(np.ones((H, W)), ux, uy)
    mask = G.periodic_mask(W, H)
    p = O.params(W, H, G.nu_tau_inv(5e-3), cs=0.0, bc=O.PERIODIC)
    h = O.deviations(f0, fmt)
    f64 = f0.copy()
    par = 0
    for _ in range(200):
        O.aa_step_h16(p, mask, h, fmt, np.zeros(H, np.float32), par)
        O.aa_step(p, mask, f64, np.zeros(H), par)
        par ^= 1
    _, ux64, uy64 = O.macro(p, mask, f64, np.zeros(H), par)
    _, uxh, uyh = O.macro(p, mask, O.populations(h, fmt), np.zeros(H), par)
    err = np.sqrt(((uxh - ux64) ** 2 + (uyh - uy64) ** 2).sum() / (ux64 ** 2 + uy64 ** 2).sum())
    assert err < tol, err


def test_h16_storage_clamp_is_the_population_clamp(oracle_lib):
    """Stored deviations stay in [-w, 10 - w], i.e. populations in [0, 10] (lbm_core4.cu:203)."""
    W, H = 32, 16
    mask = G.channel_mask(W, H)
    prof = G.parabolic_profile(H, 0.3)
    p = O.params(W, H, 1.9, cs=0.2)
    h = O.deviations(G.channel_initial_state(W, H, prof, noise=0.05, seed=3), O.F16)
    par = 0
    for _ in range(20):
        O.aa_step_h16(p, mask, h, O.F16, prof.astype(np.float32), par)
        par ^= 1
    f = O.populations(h, O.F16)
    sel = O.observable(p, mask, par)
    assert f[sel].min() >= -1e-6 and f[sel].max() <= 10 + 1e-6
