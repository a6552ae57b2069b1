"""The learned-closure step on the GPU: the neq pre-pass equals the oracle's FP32 mirror bit for bit; the
step's populations equal the oracle's AA step with the GPU's per-cell rate 1 / (tau_0 + 3 nu_t) bit for bit;
nu_t agrees with the CPU FP64 MLP (BF16 mirror) on the same features; with every cell outside the trust
region the step is exactly the Smagorinsky step."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from oracle import mlp as M
from lbm import geometry as G

pytestmark = pytest.mark.gpu


def _sim(W=200, H=96, seed=3, cs=0.1):
    from lbm.sim import Simulation
    ux, uy = G.random_phase_velocity(W, H, k0=6, u_rms=0.05, seed=seed)
    mask = G.periodic_mask(W, H)
    mask[40:46, 60:69] = 1
    return Simulation(W, H, 1.9, cs=cs, model="mrt", bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy))


def _op(sim):
    p = sim.params
    return O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp)


def _model(sim, seed=8, nu_scale=2e-3):
    from lbm.closure import MLP, ClosureConfig, neq
    n = neq(sim).cpu().numpy()
    scale = 1.0 / np.maximum(n.reshape(2, -1).std(1), 1e-12)
    cfg = ClosureConfig(in_scale=np.tile(scale, 9).astype(np.float32), nu_scale=nu_scale, nu_max=0.05, feat_max=8.0)
    p = M.init(seed)
    p[M.OFF_W3:M.OFF_B3] *= 0.2
    p[M.OFF_B3] = 3.0  # outputs around 3: mostly inside [0, nu_max / nu_scale]
    return MLP(p), cfg, p


@pytest.mark.parametrize("parity", [0, 1])
def test_neq_bitwise(gpu, parity):
    from lbm.closure import neq
    sim = _sim()
    sim.step(3 + parity)
    got = neq(sim).cpu().numpy()
    f, par = sim.state()
    assert par == (3 + parity) & 1
    ref = O.neq(_op(sim), sim.mask_host, f, np.zeros(sim.H, np.float32), par)
    assert np.array_equal(got, ref)
    assert not got[:, sim.mask_host.astype(bool)].any()


@pytest.mark.parametrize("parity", [0, 1])
def test_closure_step_bitwise_given_nut(gpu, parity):
    from lbm.closure import ClosureSim, stencil_features
    sim = _sim(seed=5)
    sim.step(20 + parity)
    mlp, cfg, p = _model(sim)
    cl = ClosureSim(sim, mlp, cfg)
    op = _op(sim)
    zero = np.zeros(sim.H, np.float32)
    for n in range(4):
        f0, par = sim.state()
        cl.step(1, record_nut=True)
        nut = cl.nut.cpu().numpy()
        feats = cl.neq.cpu().numpy()
        # the features the step used are the oracle's neq of the state before the step
        assert np.array_equal(feats, O.neq(op, sim.mask_host, f0, zero, par)), f"step {n}: features"
        fluid = sim.mask_host == 0
        ok = nut >= 0
        assert ok[fluid].mean() > 0.5, f"most cells should take the learned rate ({ok[fluid].mean():.3f})"
        # populations: oracle AA step with the GPU's per-cell rate
        tau0 = np.float32(1) / np.float32(sim.params.tau_inv)
        w_cell = np.where(ok, np.float32(1) / (tau0 + np.float32(3) * nut.astype(np.float32)), np.float32(-1))
        ref = f0.copy()
        O.aa_step_rate(op, sim.mask_host, ref, zero, par, w_cell.astype(np.float32))
        got, par2 = sim.state()
        assert par2 == par ^ 1
        sel = O.observable(op, sim.mask_host, par2)
        assert np.array_equal(got[sel], ref[sel]), f"step {n}: populations"
        # nu_t against the CPU FP64 MLP on the same scaled features
        x = stencil_features(feats) * cfg.in_scale[None, :]
        y, _ = M.forward(p, x.astype(np.float32), mirror=True)
        raw_cpu = (y * cfg.nu_scale).reshape(sim.H, sim.W)
        nu_cpu = np.maximum(raw_cpu, 0.0)  # negative outputs clamp to 0
        both = ok & fluid & (raw_cpu <= cfg.nu_max)
        err = np.abs(nut[both] - nu_cpu[both])
        assert err.max() <= 2e-3 * np.abs(nu_cpu[both]).max() + 1e-9, err.max()
        # the trust region: a cell takes the learned rate iff its features are in range and nu_t <= nu_max
        trusted = np.all(np.abs(x) <= cfg.feat_max, axis=1).reshape(sim.H, sim.W)
        margin = 1e-3 * cfg.nu_max
        clear = fluid & (np.abs(raw_cpu - cfg.nu_max) > margin)
        want = trusted & (raw_cpu <= cfg.nu_max)
        assert np.array_equal(ok[clear], want[clear]), f"step {n}: trust region"


def test_closure_fallback_is_smagorinsky(gpu):
    """feat_max = 0 puts every cell with a non-zero feature outside the trust region: the step is then the
    Smagorinsky step of lbm_step_aa, bit for bit, and the fallback counter counts every fluid cell."""
    from lbm.closure import ClosureSim
    sim = _sim(seed=9)
    ref = _sim(seed=9)
    sim.step(1)
    ref.step(1)
    mlp, cfg, _ = _model(sim)
    cfg.feat_max = 0.0
    cl = ClosureSim(sim, mlp, cfg)
    cl.step(2, record_nut=True)
    ref.step(2)
    assert torch.equal(sim.f, ref.f)
    assert int(cl.fallbacks.item()) == 2 * int((sim.mask_host == 0).sum())
    assert np.all(cl.nut.cpu().numpy()[sim.mask_host == 0] == -1.0)  # solid cells are not updated


def test_closure_argument_errors(gpu):
    from lbm import _lib
    from lbm.closure import ClosureConfig
    sim = _sim()
    L = _lib.load()
    c = ClosureConfig(nu_max=-1.0).c_struct()
    rc = L.lbm_step_closure(ctypes.byref(sim.params), ctypes.c_void_p(sim.mask.data_ptr()),
                            ctypes.c_void_p(sim.f.data_ptr()), 0, ctypes.c_void_p(1), ctypes.c_void_p(1),
                            ctypes.byref(c), ctypes.c_void_p(1), None, None, None)
    assert rc == -1


def test_study_pipeline_small(gpu):
    """The whole learned-closure study at 512^2 -> 64^2: data, training on the GPU kernels, a-priori and
    a-posteriori comparisons all produce finite numbers, and the network beats the zero predictor a priori."""
    from lbm.closure import StudyConfig, run_study
    cfg = StudyConfig(N=512, factor=8, spinup=200, snapshots=3, interval=100, eval_steps=40, eval_every=20,
                      epochs=3, batch=1024, lr=3e-3, window=5, augment=False)
    rep, mlp, ccfg = run_study(cfg, log=lambda *a: None)
    assert rep["train_steps"] > 0 and rep["samples_val"] == 64 * 64
    ap = rep["apriori"]
    assert all(np.isfinite(v) for v in ap.values())
    assert ap["val_wmse_mlp"] < ap["val_wmse_zero"]
    print(f"\nheld-out weighted MSE: untrained {ap['val_wmse_untrained']:.4f} -> trained {ap['val_wmse_mlp']:.4f}; "
          f"Smagorinsky {ap['val_wmse_smagorinsky']:.4f}, zero {ap['val_wmse_zero']:.4f}")
    assert ap["val_wmse_mlp"] < ap["val_wmse_untrained"]  # what training on the GPU kernels gains
    for row in rep["aposteriori"]:
        assert all(np.isfinite(v) for v in row.values())
    assert 0.0 <= rep["fallback_fraction"] <= 1.0


def test_learned_closure_beats_smagorinsky(gpu):
    """The study at its own size: fine 4096^2 MRT + Smagorinsky run (Cs 0.05), 8x box filter, the network trained
    on the GPU kernels (D4-augmented stencils, 33 x 33 least-squares eddy-viscosity targets), then 20000 coarse
    512^2 steps with the learned closure and with plain Smagorinsky (Cs 0.1, lbm_core4.cu:171-177) from the same
    filtered state, compared with the filtered fine run every 2000 coarse steps.  The learned closure must beat
    Smagorinsky on the time-averaged energy spectrum (mean |log <E>/<E_ref>| over shells 1..64) and on the
    time-averaged enstrophy, and a priori (correlation with the held-out targets and weighted squared error).
    Seed 5 (deterministic: every kernel and the training are bitwise reproducible); the same recipe also won on
    seeds 6 and 7 when it was chosen."""
    from lbm.closure import StudyConfig, run_study
    rep, _, _ = run_study(StudyConfig(), log=lambda *a: None)
    s, ap = rep["summary"], rep["apriori"]
    print(f"\nspectrum (time-avg) learned {s['spectrum_err_timeavg_learned']:.4f} vs Smagorinsky "
          f"{s['spectrum_err_timeavg_smagorinsky']:.4f}; enstrophy (time-avg) learned "
          f"{s['enstrophy_err_timeavg_learned']:.4f} vs {s['enstrophy_err_timeavg_smagorinsky']:.4f}; a priori corr "
          f"{ap['corr_mlp']:.3f} vs {ap['corr_smagorinsky']:.3f}, wmse {ap['val_wmse_mlp']:.3f} vs "
          f"{ap['val_wmse_smagorinsky']:.3f}")
    assert rep["config"]["eval_steps"] == 20000 and rep["coarse"] == 512
    assert s["spectrum_err_timeavg_learned"] < s["spectrum_err_timeavg_smagorinsky"]
    assert s["enstrophy_err_timeavg_learned"] < s["enstrophy_err_timeavg_smagorinsky"]
    assert ap["corr_mlp"] > 1.5 * ap["corr_smagorinsky"]
    assert ap["val_wmse_mlp"] < ap["val_wmse_smagorinsky"] and ap["val_wmse_mlp"] < ap["val_wmse_zero"]
    assert rep["fallback_fraction"] < 0.01
