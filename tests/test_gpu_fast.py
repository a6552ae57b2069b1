"""GPU: the fast arithmetic (lbm_params.fast = 1) -- the collision's reciprocals and square roots are the hardware's
v_rcp_f32 / v_sqrt_f32 (within 1 ulp), the arithmetic of the reference's -use_fast_math build, so no CPU mirror
reproduces its bits.  Parity is held to the FP64 oracle at the stated tolerance (1e-6 relative per population per
step over the 512 x 128 cylinder's 1000 steps) and to the exact arithmetic within a few ulps; every blocked form
stays bit-identical to fast single steps (two- and four-step tile kernels, every tile shape, channel, periodic and
narrow lattices, 16-bit and row-based storage, and the full 10 000-step 4096 x 1024 cylinder run)."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm import _lib
from lbm import configs
from lbm import geometry as G

pytestmark = pytest.mark.gpu

REL_TOL_FP64 = 1e-6


def _op(sim):
    p = sim.params
    return O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp)


def _sim(case="channel", W=300, H=72, storage="f32", row_base=None, arith="fast", cs=0.12, clamp=True, seed=6):
    from lbm.sim import Simulation
    if case == "periodic":
        mask = G.periodic_mask(W, H)
        mask[20:25, 60 % W:66 % W] = 1
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=seed)
        return Simulation(W, H, 1.7, cs=cs, bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy),
                          storage=storage, row_base=row_base, arith=arith, clamp=clamp)
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 6))
    mask[20:26, 200 % W:206 % W] = 1
    prof = G.parabolic_profile(H, 0.05)
    return Simulation(W, H, 1.75, cs=cs, mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed),
                      storage=storage, row_base=row_base, arith=arith, clamp=clamp)


def test_fast_cylinder_small_1000_steps_vs_fp64(gpu):
    """512 x 128 cylinder, 1000 fast single steps: at every step within 1e-6 relative per population of the FP64
    oracle advanced one step from the GPU's previous state."""
    from lbm.sim import Simulation
    case = configs.cylinder_small()
    sim = Simulation(case.W, case.H, case.tau_inv, arith="fast", **case.sim_kwargs())
    p = _op(sim)
    prof64 = sim.inlet_u_host.astype(np.float64)
    f_prev, par = sim.state()
    worst = 0.0
    for n in range(case.steps):
        sim.step(1)
        f_gpu, _ = sim.state()
        sel = O.observable(p, sim.mask_host, par ^ 1)
        r64 = f_prev.astype(np.float64)
        O.aa_step(p, sim.mask_host, r64, prof64, par)
        rel = np.abs(f_gpu[sel] - r64[sel]) / np.abs(r64[sel])
        worst = max(worst, float(rel.max()))
        assert rel.max() <= REL_TOL_FP64, f"step {n}: {rel.max():.3e}"
        f_prev, par = f_gpu, par ^ 1
    print(f"\nfast arithmetic: per-step worst rel err vs FP64 {worst:.3e}")


@pytest.mark.parametrize("case", ["channel", "periodic"])
def test_fast_close_to_exact(gpu, case):
    """From the same state, one fast and one exact step differ by at most a few ulps per population; over 200 steps
    the velocity fields stay within 1e-5 relative L2 of each other."""
    a = _sim(case, arith="fast")
    b = _sim(case, arith="exact")
    p = _op(a)
    for n in range(200):
        a.step(1)
        b.step(1)
        if n == 0:
            fa, fb = a.state()[0], b.state()[0]
            sel = O.observable(p, a.mask_host, a.parity)
            ulps = np.abs(fa[sel].view(np.int32).astype(np.int64) - fb[sel].view(np.int32).astype(np.int64))
            assert ulps.max() <= 8, ulps.max()
            # near rho = 1 the hardware estimates mostly round as the exact forms do: few or no values differ
            print(f"\none step: {np.count_nonzero(ulps)} of {ulps.size} values differ, at most {ulps.max()} ulp")
    _, ua, va = (t.cpu().numpy().astype(np.float64) for t in a.macro())
    _, ub, vb = (t.cpu().numpy().astype(np.float64) for t in b.macro())
    upd = O.observable(p, a.mask_host, 1)[0]
    err = np.sqrt((((ua - ub) ** 2 + (va - vb) ** 2)[upd]).sum() / ((ub ** 2 + vb ** 2)[upd]).sum())
    assert err < 1e-5, err


def _equal(a, b, sel):
    n = int(((a.f != b.f) & sel).sum())
    assert n == 0, f"{n} differ"


@pytest.mark.parametrize("case,W,H", [("channel", 301, 90), ("channel", 140, 33), ("periodic", 300, 72),
                                      ("periodic", 97, 41)])
@pytest.mark.parametrize("clamp", [True, False])
def test_fast_tile_kernels_equal_single(gpu, case, W, H, clamp):
    for k, chunks in ((2, (8, 12, 16, 24)), (4, (8, 12, 14, 16, 20, 24, 32))):
        for chunk in chunks:
            a = _sim(case, W=W, H=H, clamp=clamp)
            b = _sim(case, W=W, H=H, clamp=clamp)
            sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
            for _ in range(2):
                a.step(k * 2)
                b.step(k * 2, k=k, cfg=(0, chunk))
                assert a.parity == b.parity == 0
                _equal(a, b, sel)


@pytest.mark.parametrize("storage,row_base", [("f16", None), ("bf16", None), ("f16", "inlet"), ("bf16", "inlet")])
def test_fast_h16_pair_equals_single(gpu, storage, row_base):
    for case, W, H in (("channel", 301, 90), ("periodic", 300, 72)):
        rb = row_base if case == "channel" else None
        a = _sim(case, W=W, H=H, storage=storage, row_base=rb)
        b = _sim(case, W=W, H=H, storage=storage, row_base=rb)
        sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
        for chunk in (8, 12, 16, 24):
            a.step(4)
            b.step(4, k=2, cfg=(0, chunk))
            _equal(a, b, sel)


def test_fast_quad_cylinder_large(gpu):
    """The 4096 x 1024 cylinder's full 10 000 steps: fast four-step launches (64 x 12 tiles) bit-identical to fast
    single steps."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    a = Simulation(case.W, case.H, case.tau_inv, arith="fast", **case.sim_kwargs())
    b = Simulation(case.W, case.H, case.tau_inv, arith="fast", **case.sim_kwargs())
    sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
    for chunk in range(5):
        a.step(2000)
        b.step(2000, k=4, cfg=(0, 12))
        _equal(a, b, sel)
    assert np.isfinite(b.diagnostics()).all()


def test_fast_pair_f16r_cylinder_large(gpu):
    """The same 10 000 steps with FP16 storage relative to the inlet profile and the 12-row two-step kernel."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    kw = dict(arith="fast", storage="f16", row_base="inlet", **case.sim_kwargs())
    a = Simulation(case.W, case.H, case.tau_inv, **kw)
    b = Simulation(case.W, case.H, case.tau_inv, **kw)
    sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
    for chunk in range(5):
        a.step(2000)
        b.step(2000, k=2, cfg=(0, 12))
        _equal(a, b, sel)


def test_fast_rejected_where_not_built(gpu):
    from lbm.sim import Simulation
    sim = _sim("channel")
    with pytest.raises(_lib.LbmError, match="fast"):
        sim.step(4, k=4, cfg=(128, 32))  # strip kernel
    with pytest.raises(_lib.LbmError, match="fast"):
        Simulation(64, 32, 1.5, arith="fast", model="bgk").step(1)
    with pytest.raises(_lib.LbmError, match="fast"):
        Simulation(64, 32, 1.5, arith="fast", sums="neumaier").step(1)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_fast_ensemble_members_equal_solo(gpu, storage):
    """An ensemble of uneven periodic lattices with the fast arithmetic: single-step, two-step and (FP32) four-step
    launches leave every member bit-identical to its solo run with fast single steps; members must agree."""
    from lbm.ensemble import Ensemble
    from lbm.sim import Simulation
    sizes = [(300, 72), (97, 41), (256, 130), (140, 33)]
    sims, solo = [], []
    for i, (W, H) in enumerate(sizes):
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=10 + i)
        f0 = G.feq(np.ones((H, W)), ux, uy)
        kw = dict(cs=0.1, bc="periodic", f0=f0, storage=storage, arith="fast")
        sims.append(Simulation(W, H, 1.6 + 0.05 * i, **kw))
        solo.append(Simulation(W, H, 1.6 + 0.05 * i, **kw))
    ens = Ensemble(sims, two_step=True)
    ens.chunk = 12
    ens.step(4, k=2)
    ens.step(3, k=1)
    ens.step(1, k=1)
    if storage == "f32":
        ens.step(8, k=4)
    n = 16 if storage == "f32" else 8
    for s in solo:
        s.step(n)
    for m, s in zip(ens.members, solo):
        sel = torch.from_numpy(O.observable(_op(s), s.mask_host, s.parity, ghost=True)).to(s.device)
        assert m.parity == s.parity
        assert int(((m.f != s.f) & sel).sum()) == 0
    ens.close()
    mixed = [Simulation(64, 32, 1.6, bc="periodic", arith="fast"), Simulation(64, 32, 1.6, bc="periodic")]
    with pytest.raises(_lib.LbmError, match="fast"):
        Ensemble(mixed)
