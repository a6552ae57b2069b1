"""GPU: the precision workload -- periodic double shear layer and Taylor-Green vortex, MRT without LES
(Cs = 0) at nu = 1e-4 (tau_inv = 1/(3 nu + 1/2) = 1.99880..., the stress moments relax almost at the
stability limit), stored as FP32 populations or FP16 / BF16 deviations.

* 512 x 128 shear layer: every single step equals the oracle's FP32 / 16-bit mirror bit for bit (50 steps),
  and the FP32 step is within 1e-6 relative per population of the FP64 oracle advanced from the GPU state.
* The two-step tile kernel (all tile heights) and the strip-sweep kernel (k = 4, 8) equal single steps bit for
  bit on the same lattice, so the les = 0 branch of every temporally blocked kernel is covered in all storages.
* 8192 x 2048 (the workload's own size): the 16-bit two-step kernel equals 16-bit single steps bit for bit
  over 2000 steps; the Taylor-Green vortex decays as exp(-nu (kx^2 + ky^2) t) over the full 20000 steps.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
from lbm import configs
from lbm import geometry as G

pytestmark = pytest.mark.gpu

FMTS = {"f32": None, "f16": O.F16, "bf16": O.BF16}
REL_TOL_FP64 = 1e-6  # FP32 single step vs FP64 oracle, per population, per step


def _shear(storage, W=512, H=128, seed=3):
    from lbm.sim import Simulation
    case = configs.shear_layer(W=W, H=H, steps=50, seed=seed)
    assert case.cs == 0.0 and case.bc == "periodic" and abs(case.tau_inv - 1.9988) < 1e-4
    return Simulation(case.W, case.H, case.tau_inv, storage=storage, **case.sim_kwargs())


def _op(sim):
    p = sim.params
    return O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp, sums=p.sums)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_shear_layer_single_step_bitwise(gpu, storage):
    """50 single steps: each equals the oracle's mirror of that storage applied to the GPU's previous state.
    The lattice is periodic without obstacles, so every cell is observable."""
    sim = _shear(storage)
    op = _op(sim)
    assert sim.params.cs == 0.0
    prev, par = sim.state()
    zero = np.zeros(sim.H, np.float32)
    for n in range(50):
        sim.step(1)
        got, gpar = sim.state()
        ref = prev.copy()
        if storage == "f32":
            O.aa_step(op, sim.mask_host, ref, zero, par)
        else:
            O.aa_step_h16(op, sim.mask_host, ref, FMTS[storage], zero, par)
        assert gpar == par ^ 1
        bad = np.count_nonzero(got != ref)
        assert bad == 0, f"{storage} step {n}: {bad} stored values differ"
        prev, par = got, gpar


def test_shear_layer_fp32_vs_fp64_per_step(gpu):
    """FP32 single step within 1e-6 relative per population of the FP64 oracle, 50 steps at tau_inv ~ 1.9988."""
    sim = _shear("f32")
    op = _op(sim)
    prev, par = sim.state()
    zero = np.zeros(sim.H)
    worst = 0.0
    for n in range(50):
        sim.step(1)
        got, gpar = sim.state()
        r64 = prev.astype(np.float64)
        O.aa_step(op, sim.mask_host, r64, zero, par)
        rel = float((np.abs(got - r64) / np.abs(r64)).max())
        worst = max(worst, rel)
        assert rel <= REL_TOL_FP64, f"step {n}: {rel:.3e}"
        prev, par = got, gpar
    print(f"\nshear layer 512x128: worst per-step FP32 vs FP64 relative error {worst:.3e}")


def _equal_launches(storage, k, cfg, launches, W=512, H=128):
    a, b = _shear(storage, W, H), _shear(storage, W, H)
    for launch in range(launches):
        a.step(k)
        b.step(k, k=k, cfg=cfg)
        assert a.parity == b.parity
        n = int((a.f != b.f).sum())
        assert n == 0, f"{storage} k={k} cfg={cfg} launch {launch}: {n} stored values differ"


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("rows", [8, 12, 16, 24])
def test_shear_layer_two_step_kernel(gpu, storage, rows):
    _equal_launches(storage, 2, (0, rows), 25)


@pytest.mark.parametrize("rows", [8, 12, 16, 20, 24, 32])
def test_shear_layer_four_step_kernel(gpu, rows):
    _equal_launches("f32", 4, (0, rows), 12)


@pytest.mark.parametrize("storage,k,cfg", [("f32", 4, (64, 32)), ("f32", 8, (128, 64)), ("f32", 3, (64, 128)),
                                           ("f16", 4, (128, 64)), ("bf16", 8, (128, 32)), ("f16", 5, (128, 128))])
def test_shear_layer_strip_kernel(gpu, storage, k, cfg):
    _equal_launches(storage, k, cfg, 8)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_shear_layer_full_size_two_step(gpu, storage):
    """8192 x 2048 shear layer: the 12-row two-step kernel equals single steps bit for bit over 2000 steps
    (compared on the GPU every 100 launches and at the end)."""
    from lbm.sim import Simulation
    case = configs.shear_layer()
    a = Simulation(case.W, case.H, case.tau_inv, storage=storage, **case.sim_kwargs())
    b = Simulation(case.W, case.H, case.tau_inv, storage=storage, **case.sim_kwargs())
    del case
    for launch in range(1000):
        a.step(2)
        b.step(2, k=2, cfg=(0, 12))
        if launch % 100 == 99:
            n = int((a.f != b.f).sum())
            assert n == 0, f"{storage} launch {launch}: {n} stored values differ"
    rho, ux, uy = (t.double() for t in b.macro())
    assert torch.isfinite(rho).all() and torch.isfinite(ux).all()
    assert abs(float(rho.mean()) - 1.0) < 1e-4


def _tgv_amplitude(sim, ux0, uy0, norm):
    _, ux, uy = sim.macro()
    return float(((ux.double() * ux0).sum() + (uy.double() * uy0).sum()) / norm), ux, uy


def _tgv_run(mx, my, marks, storages=("f32", "f16", "bf16")):
    """8192 x 2048 Taylor-Green vortex array of (mx, my) periods (u0 = 0.04, consistent initial pressure,
    nu = 1e-4): FP32 with single steps, 16-bit storage with the 12-row two-step kernel.  Returns
    {(storage, steps): (amplitude, analytic decay, shape residual)} and nu_eff = (1/tau_inv - 1/2)/3 at the
    FP32 rate the kernels use.  The amplitude is the projection of u on the initial mode; the shape residual is
    |u - amplitude * u_mode| / |amplitude * u_mode|."""
    from lbm.sim import Simulation
    W, H, u0, nu = 8192, 2048, 0.04, 1e-4
    ux, uy = G.taylor_green_velocity(W, H, u0, mx=mx, my=my)
    f0 = G.feq(G.taylor_green_density(W, H, u0, mx=mx, my=my), ux, uy)
    tau_inv = float(np.float32(G.nu_tau_inv(nu)))
    nu_eff = (1.0 / tau_inv - 0.5) / 3.0
    k2 = (2 * math.pi * mx / W) ** 2 + (2 * math.pi * my / H) ** 2
    dev = torch.device("cuda")
    ux0 = torch.from_numpy(ux).to(dev)
    uy0 = torch.from_numpy(uy).to(dev)
    norm = float((ux0 * ux0 + uy0 * uy0).sum())
    del ux, uy
    out = {}
    for storage in storages:
        sim = Simulation(W, H, tau_inv, cs=0.0, model="mrt", bc="periodic", mask=G.periodic_mask(W, H), f0=f0,
                         storage=storage)
        done = 0
        for mark in marks:
            if storage == "f32":
                sim.step(mark - done)
            else:
                sim.step(mark - done, k=2, cfg=(0, 12))
            done = mark
            amp, ux, uy = _tgv_amplitude(sim, ux0, uy0, norm)
            res = float(((ux.double() - amp * ux0) ** 2 + (uy.double() - amp * uy0) ** 2).sum() / (amp * amp * norm))
            out[(storage, mark)] = (amp, math.exp(-nu_eff * k2 * mark), math.sqrt(res))
        del sim
        torch.cuda.empty_cache()
    return out, nu_eff, k2


def test_taylor_green_full_size_decay(gpu):
    """The workload's vortex (one period across 8192 x 2048), 20000 steps.  Its viscous decay over the run is
    only 2e-5, so this pins that nothing else changes the amplitude: FP32 within 2e-3 of the analytic decay (the
    lattice's start-up and compressibility error is ~6e-4), FP16 within 5e-3 and BF16 within 3e-2 of it at 5000,
    10000 and 20000 steps.  The field keeps the mode's shape through 10000 steps (residual after the projection
    below 1e-2 for FP32 -- measured ~5e-4 --, 2e-2 for FP16, whose storage rounding leaves ~1e-2 in the lowest
    modes, and 5e-2 for BF16).  Later the FP32 residual grows from a grid-scale mode (wavelength ~2.8 cells in
    x): the MRT scheme with the reference's rates (s = 1.1) is linearly unstable for uniform flow near
    u = 0.04 at tau_inv = 1.9988, which the FP64 oracle shows too (test_oracle_properties), so past 10000 steps
    the shape is reported, not gated."""
    rep, _, _ = _tgv_run(1, 1, (5000, 10000, 20000))
    print("\nTaylor-Green (1,1) 8192x2048 (amplitude error, shape residual):",
          {f"{s}@{m}": (f"{abs(a / w - 1):.2e}", f"{r:.2e}") for (s, m), (a, w, r) in rep.items()})
    tol = {"f32": 2e-3, "f16": 5e-3, "bf16": 3e-2}
    shape_tol = {"f32": 1e-2, "f16": 2e-2, "bf16": 5e-2}
    for (st, mark), (amp, want, res) in rep.items():
        assert abs(amp / want - 1.0) <= tol[st], f"{st} step {mark}: amplitude {amp:.6f} vs {want:.6f}"
        if mark <= 10000:
            assert res <= shape_tol[st], f"{st} step {mark}: shape residual {res:.2e}"


def test_taylor_green_viscous_decay(gpu):
    """A vortex array of 256 x 64 periods on the same 8192 x 2048 lattice (32 cells per wavelength, peak speed
    0.04/sqrt(2) = 0.028), where the viscous decay is measurable: exp(-nu k^2 t) = 0.891 over 15000 steps.
    FP32: the decay rate between 5000 and 20000 steps (after the start-up transient) gives nu = (1/tau_inv - 1/2)/3
    within 1 %, and the field keeps the mode's shape (residual below 1e-2).  16-bit deviation storage with
    round-to-nearest cannot follow this decay: a deviation of ~1e-2 changes by ~1e-7 per step, far below half its
    FP16 rounding step (3.8e-6), so the stored state stops changing; the test pins that property of the storage
    (the amplitude over 5000 -> 20000 stays within 2e-3 of constant), which the design notes discuss."""
    rep, nu_eff, k2 = _tgv_run(256, 64, (5000, 20000))
    nu_rate = {st: math.log(rep[(st, 5000)][0] / rep[(st, 20000)][0]) / (k2 * 15000) for st in ("f32", "f16", "bf16")}
    print(f"\nTaylor-Green (256,64): nu {nu_eff:.6e}; from the 5000->20000 decay", {k: f"{v:.4e}" for k, v in nu_rate.items()},
          "shape", {f"{s}@{m}": f"{r:.1e}" for (s, m), (a, w, r) in rep.items()})
    assert abs(nu_rate["f32"] / nu_eff - 1.0) <= 1e-2, f"FP32: nu {nu_rate['f32']:.4e} vs {nu_eff:.4e}"
    assert rep[("f32", 20000)][2] <= 1e-2
    for st in ("f16", "bf16"):
        assert abs(rep[(st, 20000)][0] / rep[(st, 5000)][0] - 1.0) <= 2e-3, f"{st}: the 16-bit state decays"
