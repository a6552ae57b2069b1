"""CPU pins of the oracle (oracle/): closed forms, an independent numpy transcription, brute-force
equivalence of the streaming forms, and analytic flows.  No GPU needed.

Each check is chosen so that a plausible transcription slip (a dropped term, a wrong sign, a swapped
index or population) fails at least one of them; none re-types the oracle's own formula.
"""
import math

import numpy as np
import pytest

import oracle as O
from oracle import pyref
from lbm import geometry as G

F32 = lambda v: float(np.float32(v))  # noqa: E731  the C oracle receives FP32 parameters


# ----------------------------------------------------------------------------------------------------
# lattice and MRT transform
# ----------------------------------------------------------------------------------------------------

def test_velocity_set_closed_forms():
    w, cx, cy = O.W_EXACT, O.CX, O.CY
    assert abs(w.sum() - 1.0) < 1e-15
    assert abs((w * cx).sum()) < 1e-15 and abs((w * cy).sum()) < 1e-15
    assert abs((w * cx * cx).sum() - 1 / 3) < 1e-15
    assert abs((w * cy * cy).sum() - 1 / 3) < 1e-15
    assert abs((w * cx * cy).sum()) < 1e-15
    assert np.all(cx[O.OPP] == -cx) and np.all(cy[O.OPP] == -cy)


# Lallemand & Luo, Phys. Rev. E 61 (2000) 6546, eq. (7): rows rho, e, eps, jx, qx, jy, qy, pxx, pxy
LALLEMAND_LUO_M = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-4, -1, -1, -1, -1, 2, 2, 2, 2],
    [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1],
    [0, -2, 0, 2, 0, 1, -1, -1, 1],
    [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1],
    [0, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1, -1, 1, -1],
], dtype=np.float64)


def test_mrt_forward_is_lallemand_luo_basis(oracle_lib):
    M = np.array([O.mrt_forward(e) for e in np.eye(9)]).T
    assert np.array_equal(M, LALLEMAND_LUO_M)


def test_mrt_inverse_is_exact(oracle_lib):
    M = np.array([O.mrt_forward(e) for e in np.eye(9)]).T
    Mi = np.array([O.mrt_inverse(e) for e in np.eye(9)]).T
    assert np.abs(M @ Mi - np.eye(9)).max() < 1e-15
    # orthogonal rows: M^-1 = M^T diag(1/|row|^2) (norms 9, 36, 36, 6, 12, 6, 12, 4, 4)
    norms = (LALLEMAND_LUO_M ** 2).sum(axis=1)
    assert np.array_equal(norms, [9, 36, 36, 6, 12, 6, 12, 4, 4])
    assert np.abs(Mi - LALLEMAND_LUO_M.T / norms).max() < 1e-16


def test_equilibrium_moments_closed_form(oracle_lib):
    rng = np.random.default_rng(0)
    for _ in range(20):
        rho = rng.uniform(0.8, 1.2)
        ux, uy = rng.uniform(-0.1, 0.1, 2)
        m = O.mrt_forward(O.feq(rho, ux, uy))
        u2 = ux * ux + uy * uy
        want = [rho, rho * (3 * u2 - 2), rho * (1 - 3 * u2), rho * ux, -rho * ux, rho * uy, -rho * uy,
                rho * (ux * ux - uy * uy), rho * ux * uy]
        assert np.abs(m - want).max() < 1e-15


# ----------------------------------------------------------------------------------------------------
# collision
# ----------------------------------------------------------------------------------------------------

def _near_eq(rng, amp=1e-3):
    rho = rng.uniform(0.95, 1.05)
    ux, uy = rng.uniform(-0.08, 0.08, 2)
    return O.feq(rho, ux, uy) + rng.uniform(-amp, amp, 9)


@pytest.mark.parametrize("model,cs", [(O.MRT, 0.0), (O.MRT, 0.15), (O.BGK, 0.0)])
def test_equilibrium_is_fixed_point(oracle_lib, model, cs):
    rng = np.random.default_rng(1)
    p = O.params(8, 8, 1.3, cs=cs, model=model)
    for _ in range(20):
        rho = rng.uniform(0.9, 1.1)
        ux, uy = rng.uniform(-0.1, 0.1, 2)
        f = O.feq(rho, ux, uy)
        g, mac = O.collide(p, f.copy())
        assert np.abs(g - f).max() < 1e-15
        assert abs(mac[0] - rho) < 1e-15 and abs(mac[1] - ux) < 1e-15 and abs(mac[2] - uy) < 1e-15


def test_mrt_with_equal_rates_is_bgk(oracle_lib):
    """All MRT rates = omega reduces to BGK (links lbm_core4's collision to lbm_core_smemaa's)."""
    rng = np.random.default_rng(2)
    for omega in (0.7, 1.3, 1.9):
        w = F32(omega)
        pm = O.params(8, 8, w, cs=0.0, model=O.MRT, clamp=0, s_e=w, s_eps=w, s_q=w)
        pb = O.params(8, 8, w, model=O.BGK)
        for _ in range(20):
            f = _near_eq(rng)
            gm, _ = O.collide(pm, f.copy())
            gb, _ = O.collide(pb, f.copy())
            assert np.abs(gm - gb).max() < 2e-15


def test_mrt_conserves_mass_and_momentum(oracle_lib):
    rng = np.random.default_rng(3)
    p = O.params(8, 8, 1.7, cs=0.2)
    for _ in range(50):
        f = _near_eq(rng, 5e-3)
        g, _ = O.collide(p, f.copy())
        assert abs(g.sum() - f.sum()) < 1e-14
        assert abs((g * O.CX).sum() - (f * O.CX).sum()) < 1e-14
        assert abs((g * O.CY).sum() - (f * O.CY).sum()) < 1e-14


def test_smagorinsky_tau_eff_solves_quadratic(oracle_lib):
    """tau_eff of lbm_core4.cu:174-177 is the positive root of tau^2 - tau0 tau - 4.5 Cs^2 Q / rho = 0,
    recovered here from how much the stress moments actually relaxed."""
    rng = np.random.default_rng(4)
    tau_inv, cs = F32(1.85), F32(0.2)
    p = O.params(8, 8, tau_inv, cs=cs)
    p0 = O.params(8, 8, tau_inv, cs=0.0)
    tau0 = 1.0 / tau_inv
    for _ in range(20):
        f = _near_eq(rng, 2e-2)
        g, mac = O.collide(p, f.copy())
        assert np.all(g > 0) and np.all(g < 10)
        rho, ux, uy = mac
        m, mp = O.mrt_forward(f), O.mrt_forward(g)
        n7 = m[7] - rho * (ux * ux - uy * uy)
        n8 = m[8] - rho * ux * uy
        omega = (m[7] - mp[7]) / n7
        assert abs((m[8] - mp[8]) / n8 - omega) < 1e-9
        Q = math.sqrt(n7 * n7 + n8 * n8 + 1e-8)
        tau = 1.0 / omega
        assert abs(tau * tau - tau0 * tau - 4.5 * cs * cs * Q / rho) < 1e-12
        assert tau > tau0
        g0, _ = O.collide(p0, f.copy())
        mp0 = O.mrt_forward(g0)
        assert abs((m[7] - mp0[7]) / n7 - tau_inv) < 1e-9


def test_zou_he_inlet_moments(oracle_lib):
    """Zou & He (1997): the rebuilt f1, f5, f8 give rho u_x = rho u and rho u_y = 0, with
    rho (1 - u) = f0 + f2 + f4 + 2 (f3 + f6 + f7).  MRT keeps those moments through the collision."""
    rng = np.random.default_rng(5)
    p = O.params(8, 8, 1.6, cs=0.2)
    for _ in range(20):
        f = _near_eq(rng)
        u = rng.uniform(0.0, 0.1)
        g, mac = O.collide(p, f.copy(), inlet=True, uin=u)
        rho = (f[0] + f[2] + f[4] + 2 * (f[3] + f[6] + f[7])) / (1 - u)
        assert abs(mac[0] - rho) < 1e-15 and mac[1] == u and mac[2] == 0.0
        assert abs(g.sum() - rho) < 1e-14
        assert abs((g * O.CX).sum() - rho * u) < 1e-14
        assert abs((g * O.CY).sum()) < 1e-14


def test_clamps_engage_as_reference(oracle_lib):
    """Velocity clamp +-0.4 and density clamp [0, 1.5] (lbm_core4.cu:142-144)."""
    p = O.params(8, 8, 1.5, cs=0.1)
    f = O.feq(1.0, 0.6, -0.5)  # |u| beyond 0.4 -> clamped
    _, mac = O.collide(p, f.copy())
    assert mac[1] == pytest.approx(0.4) and mac[2] == pytest.approx(-0.4)
    f = O.feq(2.0, 0.01, 0.0)
    g, mac = O.collide(p, f.copy())
    assert mac[0] == 1.5
    assert np.all((g >= 0) & (g <= 10))


# ----------------------------------------------------------------------------------------------------
# streaming: C oracle vs the independent numpy transcription, and AA vs pull
# ----------------------------------------------------------------------------------------------------

def _channel_case(W=40, H=18, noise=1e-3, u=0.05, seed=1):
    mask = G.channel_mask(W, H, cylinder=(12, H // 2, 3))
    prof = G.parabolic_profile(H, u)
    f0 = G.channel_initial_state(W, H, prof, noise=noise, seed=seed)
    return mask, prof, f0


@pytest.mark.parametrize("model,cs", [(O.MRT, 0.15), (O.MRT, 0.0), (O.BGK, 0.0)])
def test_c_pull_matches_numpy_transcription(oracle_lib, model, cs):
    mask, prof, f0 = _channel_case()
    W, H = mask.shape[1], mask.shape[0]
    p = O.params(W, H, 1.7, cs=cs, model=model)
    a, b = f0.copy(), f0.copy()
    for _ in range(25):
        a2 = a.copy()
        rho, ux, uy = O.pull_step(p, mask, a, a2, prof)
        a = a2
        if model == O.BGK:
            b = pyref.pull_step_bgk(b, mask, prof, F32(1.7))
        else:
            b, rb, uxb, uyb = pyref.pull_step(b, mask, prof, F32(1.7), F32(cs), F32(1.1), F32(1.1), F32(1.1))
    sel = O.observable(p, mask, 1)
    assert np.abs(a - b)[sel].max() < 1e-13
    if model != O.BGK:
        upd = sel[0]
        assert np.abs(rho - rb)[upd].max() < 1e-13 and np.abs(ux - uxb)[upd].max() < 1e-13


def test_c_pull_matches_numpy_periodic_with_obstacle(oracle_lib):
    W, H = 24, 20
    mask = G.periodic_mask(W, H)
    mask[8:11, 5:8] = 1
    ux, uy = G.taylor_green_velocity(W, H, 0.05)
    f0 = G.feq(np.ones((H, W)), ux, uy)
    p = O.params(W, H, 1.4, cs=0.1, bc=O.PERIODIC)
    zero = np.zeros(H)
    a, b = f0.copy(), f0.copy()
    for _ in range(20):
        a2 = a.copy()
        O.pull_step(p, mask, a, a2, zero)
        a = a2
        b, _, _, _ = pyref.pull_step(b, mask, zero, F32(1.4), F32(0.1), F32(1.1), F32(1.1), F32(1.1), periodic=True)
    sel = O.observable(p, mask, 1)
    assert np.abs(a - b)[sel].max() < 1e-13


def test_clamped_flow_matches_numpy(oracle_lib):
    """Clamps active on many cells: C oracle and numpy transcription still agree."""
    W, H = 30, 12
    mask = G.channel_mask(W, H)
    prof = G.parabolic_profile(H, 0.3)
    rng = np.random.default_rng(9)
    f0 = G.feq(np.ones((H, W)) * 1.3, 0.45 + 0.05 * rng.standard_normal((H, W)), 0.05 * rng.standard_normal((H, W)))
    p = O.params(W, H, 1.9, cs=0.2)
    a, b = f0.copy(), f0.copy()
    for _ in range(6):
        a2 = a.copy()
        O.pull_step(p, mask, a, a2, prof)
        a = a2
        b, _, _, _ = pyref.pull_step(b, mask, prof, F32(1.9), F32(0.2), F32(1.1), F32(1.1), F32(1.1))
    sel = O.observable(p, mask, 1)
    assert np.abs(a - b)[sel].max() < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("model,cs,bc", [(O.MRT, 0.15, O.CHANNEL), (O.BGK, 0.0, O.CHANNEL),
                                         (O.MRT, 0.1, O.PERIODIC)])
def test_aa_equals_pull_bitwise(oracle_lib, dtype, model, cs, bc):
    """AA in place vs two-buffer pull.  With P_0 = K(C_0) (one A step), the pull state P_n equals the
    post-collision state K(C_n) that the AA form holds after every A step -- bit for bit, in both
    precisions (same collision, different streaming code)."""
    if bc == O.CHANNEL:
        mask, prof, f0 = _channel_case()
    else:
        W, H = 26, 16
        mask = G.periodic_mask(W, H)
        mask[5:8, 9:12] = 1
        ux, uy = G.taylor_green_velocity(W, H, 0.05)
        f0 = G.feq(np.ones((H, W)), ux, uy)
        prof = np.zeros(H)
    H, W = mask.shape
    p = O.params(W, H, 1.7, cs=cs, model=model, bc=bc)
    prof = prof.astype(dtype)
    C = f0.astype(dtype)
    O.aa_step(p, mask, C, prof, 0)
    P = O.reverse(C).copy()
    par = 1
    sel = O.observable(p, mask, 1)
    for n in range(1, 17):
        O.aa_step(p, mask, C, prof, par)
        par ^= 1
        P2 = P.copy()
        O.pull_step(p, mask, P, P2, prof)
        P = P2
        if par == 1:
            assert np.array_equal(O.reverse(C)[sel], P[sel]), f"step {n}"


def test_outlet_ghost_is_copy_of_last_column(oracle_lib):
    mask, prof, f0 = _channel_case()
    H, W = mask.shape
    p = O.params(W, H, 1.7, cs=0.15)
    a = f0.copy()
    b = a.copy()
    O.pull_step(p, mask, a, b, prof)
    assert np.array_equal(b[:, 1:H - 1, W - 1], b[:, 1:H - 1, W - 2])
    C = f0.copy()
    O.aa_step(p, mask, C, prof, 0)  # reversed copy
    assert np.array_equal(C[:, 1:H - 1, W - 1], C[:, 1:H - 1, W - 2])


# ----------------------------------------------------------------------------------------------------
# conservation and analytic flows
# ----------------------------------------------------------------------------------------------------

def test_periodic_mass_momentum_conservation(oracle_lib):
    W, H = 32, 24
    mask = G.periodic_mask(W, H)
    ux, uy = G.random_phase_velocity(W, H, k0=3, u_rms=0.03, seed=7)
    f = G.feq(np.ones((H, W)), ux, uy)
    p = O.params(W, H, 1.6, cs=0.15, bc=O.PERIODIC)
    m0 = f.sum()
    jx0 = (f * O.CX[:, None, None]).sum()
    par = 0
    for _ in range(40):
        O.aa_step(p, mask, f, np.zeros(H), par)
        par ^= 1
    g = O.reverse(f) if par else f
    assert abs(g.sum() - m0) < 1e-11
    assert abs((g * O.CX[:, None, None]).sum() - jx0) < 1e-12


def test_bounce_back_conserves_mass_in_closed_box(oracle_lib):
    W, H = 30, 20
    mask = G.periodic_mask(W, H)
    mask[0, :] = 1
    mask[H - 1, :] = 1
    mask[:, 0] = 1
    mask[:, W - 1] = 1
    mask[8:12, 10:14] = 1
    ux, uy = G.taylor_green_velocity(W, H, 0.04)
    f = G.feq(np.ones((H, W)), ux, uy)
    p = O.params(W, H, 1.5, cs=0.0, bc=O.PERIODIC)
    fluid = ~mask.astype(bool)
    m0 = f[:, fluid].sum()
    par = 0
    for _ in range(30):
        O.aa_step(p, mask, f, np.zeros(H), par)
        par ^= 1
    assert abs(f[:, fluid].sum() - m0) < 1e-11


@pytest.mark.parametrize("model,nu", [(O.MRT, 0.1), (O.MRT, 0.02), (O.BGK, 0.05)])
def test_shear_wave_decay_between_half_way_walls(oracle_lib, model, nu):
    """u = U sin(pi (y - 1/2) / L) between no-slip walls at y = 1/2 and y = L + 1/2 decays as
    exp(-nu (pi/L)^2 t).  Half-way bounce-back puts the walls there; a half-cell shift would change the
    decay rate by ~3% at L = 32."""
    L = 32
    H, W = L + 2, 4
    mask = G.periodic_mask(W, H)
    mask[0, :] = 1
    mask[H - 1, :] = 1
    y = np.arange(H, dtype=np.float64)
    prof = 0.01 * np.sin(math.pi * (y - 0.5) / L)
    prof[0] = prof[H - 1] = 0.0
    ux = np.broadcast_to(prof[:, None], (H, W)).copy()
    f = G.feq(np.ones((H, W)), ux, np.zeros((H, W)))
    p = O.params(W, H, G.nu_tau_inv(nu), cs=0.0, bc=O.PERIODIC, model=model)
    n = int(0.5 * L * L / (nu * math.pi ** 2))
    par = O.aa_steps(p, mask, f, np.zeros(H), 0, n)
    _, u1, uy1 = O.macro(p, mask, f, np.zeros(H), par)
    amp = (u1[1:H - 1, 0] * prof[1:H - 1]).sum() / (prof[1:H - 1] ** 2).sum()
    assert abs(amp / math.exp(-nu * (math.pi / L) ** 2 * n) - 1) < 3e-3
    assert np.abs(uy1[1:H - 1]).max() < 1e-8


def test_channel_inlet_holds_prescribed_profile(oracle_lib):
    """Zou-He fixes ux = inlet_u(y), uy = 0 on the inlet column at every step (lbm_core4.cu:59-61)."""
    W, H = 48, 18
    mask = G.channel_mask(W, H)
    prof = G.parabolic_profile(H, 0.03)
    p = O.params(W, H, G.nu_tau_inv(0.05), cs=0.1)
    f = G.channel_initial_state(W, H, prof, noise=1e-3, seed=2)
    par = 0
    for _ in range(50):
        rho, ux, uy = O.aa_step(p, mask, f, prof, par, macros=True)
        par ^= 1
        assert np.array_equal(ux[1:H - 1, 0], prof[1:H - 1]) and not uy[1:H - 1, 0].any()


def test_taylor_green_decay(oracle_lib):
    W = H = 48
    nu = 0.02
    u0 = 0.01
    mask = G.periodic_mask(W, H)
    ux, uy = G.taylor_green_velocity(W, H, u0)
    f = G.feq(np.ones((H, W)), ux, uy)
    p = O.params(W, H, G.nu_tau_inv(nu), cs=0.0, bc=O.PERIODIC)
    steps = 400
    par = O.aa_steps(p, mask, f, np.zeros(H), 0, steps)
    _, ux1, uy1 = O.macro(p, mask, f, np.zeros(H), par)
    amp = (ux1 * ux).sum() / (ux * ux).sum()
    k2 = 2 * (2 * math.pi / W) ** 2
    assert abs(amp / math.exp(-nu * k2 * steps) - 1) < 5e-3  # O((k dx)^2) lattice error at 48 cells


def test_fp32_mirror_tracks_fp64_per_step(oracle_lib):
    """The per-step tolerance used on the GPU (1e-6 relative per population) holds for the FP32 mirror."""
    mask, prof, f0 = _channel_case(W=64, H=24)
    H, W = mask.shape
    p = O.params(W, H, 1.8, cs=0.1)
    f32 = f0.astype(np.float32)
    par = 0
    worst = 0.0
    for _ in range(60):
        f64 = f32.astype(np.float64)
        O.aa_step(p, mask, f64, prof, par)
        O.aa_step(p, mask, f32, prof.astype(np.float32), par)
        par ^= 1
        sel = O.observable(p, mask, par)
        rel = np.abs(f32[sel] - f64[sel]) / np.abs(f64[sel])
        worst = max(worst, rel.max())
    assert worst < 1e-6


def test_golden_fixture_regression(oracle_lib):
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "channel_32x16_mrt_les.npz")
    z = np.load(path, allow_pickle=False)
    mask, prof, f = z["mask"], z["inlet_u"], z["f0"].copy()
    H, W = mask.shape
    p = O.params(W, H, float(z["tau_inv"]), cs=float(z["cs"]))
    par = O.aa_steps(p, mask, f, prof, 0, int(z["steps"]))
    sel = O.observable(p, mask, par)
    assert np.abs(f - z["f_end"])[sel].max() < 1e-14
    rho, ux, uy = O.macro(p, mask, f, prof, par)
    assert np.abs(rho - z["rho"])[sel[0]].max() < 1e-14
