"""GPU: 16-bit storage relative to a per-row base state (lbm_params.row_base) -- single steps bit for bit against
the oracle's 16-bit mirror with the same table, the two-step kernel bit for bit against single steps (every tile
height, channel, periodic, narrow and compensated-sum cases, and the full 10 000-step 4096 x 1024 cylinder run),
conversions and macro fields, the errors the library reports for forms that do not take a base, and the accuracy
gain over the rest-state base on the scaled cylinder flow against the FP64 oracle.  The base window
(lbm_params.window: per-cell base velocities, re-derived by lbm_rebase_window) is held to the same bars: single
steps and re-derivations bit for bit against the mirror, the two-step kernel bit for bit against single steps
(also over the 4096 x 1024 cylinder's 10 000 steps with the wake window)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from lbm import _lib
from lbm import configs
from lbm import geometry as G

pytestmark = pytest.mark.gpu

FMTS = [("f16", O.F16), ("bf16", O.BF16)]


def _shear(W, H):
    """A periodic shear profile ux(y) and a small cross flow, the rows' base velocity for the periodic cases."""
    y = np.arange(H)
    return 0.05 * np.sin(2 * np.pi * (y + 0.5) / H), 0.004 * np.cos(2 * np.pi * y / H)


def _sim(storage, case="channel", W=300, H=72, cs=0.12, seed=6, sums="plain", model="mrt"):
    """channel: parabolic inlet, a cylinder and a block, base = the inlet profile.  periodic: obstacles and a
    random-phase field on a shear profile, base = that profile (both components).  W = 300 gives interior tiles."""
    from lbm.sim import Simulation
    if case == "periodic":
        mask = G.periodic_mask(W, H)
        mask[20:25, 60:66] = 1
        mask[40:47, 180 % W:186 % W] = 1
        ux, uy = G.random_phase_velocity(W, H, 4, 0.02, seed=seed)
        bx, by = _shear(W, H)
        f0 = G.feq(np.ones((H, W)), ux + bx[:, None], uy + by[:, None])
        return Simulation(W, H, 1.7, cs=cs, bc="periodic", mask=mask, f0=f0, storage=storage, sums=sums,
                          model=model, row_base=(bx, by))
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 6))
    mask[20:26, 200 % W:206 % W] = 1
    prof = G.parabolic_profile(H, 0.05)
    return Simulation(W, H, 1.75, cs=cs, mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed),
                      storage=storage, sums=sums, model=model, row_base="inlet")


def _op(sim):
    p = sim.params
    return O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp, sums=p.sums)


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("case,W,H", [("channel", 300, 72), ("channel", 131, 37), ("periodic", 300, 72),
                                      ("periodic", 97, 41)])
@pytest.mark.parametrize("model,cs,sums", [("mrt", 0.12, "plain"), ("mrt", 0.0, "plain"), ("bgk", 0.0, "plain"),
                                           ("mrt", 0.12, "neumaier")])
def test_rowbase_single_step_bitwise(gpu, storage, fmt, case, W, H, model, cs, sums):
    sim = _sim(storage, case, W=W, H=H, cs=cs, sums=sums, model=model)
    op = _op(sim)
    prev, par = sim.state()
    for n in range(12):
        sim.step(1)
        got, gpar = sim.state()
        ref = prev.copy()
        O.aa_step_h16(op, sim.mask_host, ref, fmt, sim.inlet_u_host, par, base=sim.row_base_host)
        sel = O.observable(op, sim.mask_host, par ^ 1, ghost=True)
        assert gpar == par ^ 1
        assert np.array_equal(got[sel], ref[sel]), f"step {n}: {np.count_nonzero(got[sel] != ref[sel])} differ"
        prev, par = got, gpar


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("case,W,H", [("channel", 301, 90), ("channel", 140, 33), ("periodic", 300, 72),
                                      ("periodic", 97, 41), ("periodic", 384, 96)])
@pytest.mark.parametrize("chunk", [8, 12, 16, 24])
def test_rowbase_pair_equals_single(gpu, storage, fmt, case, W, H, chunk):
    a = _sim(storage, case, W=W, H=H)
    b = _sim(storage, case, W=W, H=H)
    if case == "periodic" and W == 384:  # no obstacle: every interior tile runs the all-fluid odd step
        for s in (a, b):
            s.mask.zero_()
            s.mask_host[:] = 0
    op = _op(a)
    sels = [torch.from_numpy(O.observable(op, a.mask_host, q, ghost=True)).to(a.device) for q in (0, 1)]
    for launch in range(4):
        a.step(2)
        b.step(2, k=2, cfg=(0, chunk))
        assert a.parity == b.parity == 0
        n = int(((a.f != b.f) & sels[0]).sum())
        assert n == 0, f"chunk {chunk} launch {launch}: {n} differ"


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_rowbase_pair_compensated_sums(gpu, storage):
    a = _sim(storage, "channel", W=301, H=90, sums="neumaier")
    b = _sim(storage, "channel", W=301, H=90, sums="neumaier")
    op = _op(a)
    sel = torch.from_numpy(O.observable(op, a.mask_host, 0, ghost=True)).to(a.device)
    a.step(8)
    b.step(8, k=2, cfg=(0, 12))
    assert int(((a.f != b.f) & sel).sum()) == 0


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("case", ["channel", "periodic"])
def test_rowbase_conversion_and_macros(gpu, storage, fmt, case):
    sim = _sim(storage, case)
    op = _op(sim)
    base = sim.row_base_host
    for nsteps in (7, 1):  # reversed, then canonical again
        sim.step(nsteps)
        h, par = sim.state()
        f = (O.from_h16(h, fmt) + O._slot_base(base, par, h.shape)).astype(np.float32)
        f = (f + O.W_F32[:, None, None]).astype(np.float32)
        assert np.array_equal(sim.populations().cpu().numpy(), f), f"parity {par}"
        rho, ux, uy = (t.cpu().numpy() for t in sim.macro())
        r, a, b = O.macro(op, sim.mask_host, f, sim.inlet_u_host, par)
        upd = O.observable(op, sim.mask_host, 1)[0]
        assert np.array_equal(r[upd], rho[upd]) and np.array_equal(a[upd], ux[upd]) and np.array_equal(b[upd], uy[upd])


@pytest.mark.parametrize("storage,fmt", FMTS)
def test_rowbase_initial_state_conversion(gpu, storage, fmt):
    """The FP32 -> 16-bit conversion: half((f - w) - d) per slot, at either parity."""
    from lbm.sim import Simulation
    W, H = 200, 40
    prof = G.parabolic_profile(H, 0.05)
    f0 = G.channel_initial_state(W, H, prof, 1e-3, 3).astype(np.float32)
    for parity in (0, 1):
        sim = Simulation(W, H, 1.75, cs=0.1, mask=G.channel_mask(W, H), inlet_u=prof, f0=f0, storage=storage,
                         row_base="inlet", parity=parity)
        d = O._slot_base(sim.row_base_host, parity, f0.shape)
        want = O.to_h16(((f0 - O.W_F32[:, None, None]).astype(np.float32) - d).astype(np.float32), fmt)
        assert np.array_equal(sim.state()[0], want), f"parity {parity}"


def test_rowbase_checkpoint_roundtrip(gpu, tmp_path):
    from lbm.sim import Simulation
    sim = _sim("f16")
    sim.step(5)
    sim.save(tmp_path / "rb.npz")
    back = Simulation.load(tmp_path / "rb.npz")
    assert back.row_base_host is not None and np.array_equal(back.row_base_host, sim.row_base_host)
    assert back.parity == sim.parity and torch.equal(back.f, sim.f)
    sim.step(3)
    back.step(3)
    assert torch.equal(back.f, sim.f)


def test_rowbase_rejected_forms(gpu):
    """Forms without a base path say so instead of running: the strip kernel, odd-parity or k != 2 tile launches,
    lbm_convert (the base depends on the parity), ensembles."""
    sim = _sim("f16")
    with pytest.raises(_lib.LbmError, match="row-based"):
        sim.step(4, k=4, cfg=(128, 32))
    with pytest.raises(_lib.LbmError, match="row-based"):
        sim.step(2, k=2, cfg=(128, 32))
    sim.step(1)
    with pytest.raises(_lib.LbmError):
        sim.step(2, k=2, cfg=(0, 12))  # odd parity
    L = _lib.load()
    out = torch.empty(sim.f.shape, dtype=torch.float32, device=sim.device)
    rc = L.lbm_convert(ctypes.byref(sim.params), ctypes.c_void_p(sim.f.data_ptr()), _lib.FMT_F16,
                       ctypes.c_void_p(out.data_ptr()), _lib.FMT_F32, None)
    assert rc == -1 and b"lbm_convert_state" in L.lbm_last_error()


def test_rowbase_pair_cylinder_large(gpu):
    """The 4096 x 1024 cylinder run of the precision workload, 10 000 steps: the two-step kernel with FP16 storage
    relative to the inlet profile stays bit-identical to single steps."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    a = Simulation(case.W, case.H, case.tau_inv, storage="f16", row_base="inlet", **case.sim_kwargs())
    b = Simulation(case.W, case.H, case.tau_inv, storage="f16", row_base="inlet", **case.sim_kwargs())
    op = _op(a)
    sel = torch.from_numpy(O.observable(op, a.mask_host, 0, ghost=True)).to(a.device)
    for chunk in range(5):
        a.step(2000)
        b.step(2000, k=2, cfg=(0, 12))
        n = int(((a.f != b.f) & sel).sum())
        assert n == 0, f"after {2000 * (chunk + 1)} steps: {n} differ"


def test_rowbase_accuracy_gain_scaled_cylinder(gpu, oracle_lib):
    """The cylinder flow at 1024 x 256 (Re 1000 geometry scaled), 1000 steps: velocity L2 against the FP64 oracle
    with FP16 storage relative to the inlet profile is below 2.5e-4 and at most a fifth of the rest-state base's."""
    from lbm.sim import Simulation
    case = configs.cylinder_scaled(1024, 256, steps=1000)
    steps = 1000
    p = _op(Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs()))
    f64 = case.f0.astype(np.float64).copy()
    par = O.aa_steps(p, case.mask, f64, np.asarray(case.inlet_u, np.float64), 0, steps)
    R, A, B = O.macro(p, case.mask, f64, np.asarray(case.inlet_u, np.float64), par)
    upd = O.observable(p, case.mask, 1)[0]
    errs = {}
    for rb in (None, "inlet"):
        sim = Simulation(case.W, case.H, case.tau_inv, storage="f16", row_base=rb, **case.sim_kwargs())
        sim.step(steps, k=2)
        r, a, b = (t.cpu().numpy().astype(np.float64) for t in sim.macro())
        errs[rb] = np.sqrt((((a - A) ** 2 + (b - B) ** 2)[upd]).sum() / ((A ** 2 + B ** 2)[upd]).sum())
    assert errs["inlet"] < 2.5e-4, errs
    assert errs["inlet"] < errs[None] / 5, errs


# ------------------------------------------------------------------ per-cell base window
def _wsim(storage, case="channel", W=300, H=72, window=(60, 20, 90, 30), rebase_every=0, arith="exact", seed=6):
    """_sim's lattices with a base window (lbm_params.window); channel: from the cylinder's trailing side over
    the block; the window crosses tile boundaries of every chunk height."""
    from lbm.sim import Simulation
    if case == "periodic":
        mask = G.periodic_mask(W, H)
        mask[20:25, 60:66] = 1
        ux, uy = G.random_phase_velocity(W, H, 4, 0.02, seed=seed)
        bx, by = _shear(W, H)
        f0 = G.feq(np.ones((H, W)), ux + bx[:, None], uy + by[:, None])
        return Simulation(W, H, 1.7, cs=0.12, bc="periodic", mask=mask, f0=f0, storage=storage, row_base=(bx, by),
                          window=window, rebase_every=rebase_every, arith=arith)
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 6))
    mask[20:26, 200 % W:206 % W] = 1
    prof = G.parabolic_profile(H, 0.05)
    return Simulation(W, H, 1.75, cs=0.12, mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed),
                      storage=storage, row_base="inlet", window=window, rebase_every=rebase_every, arith=arith)


WINDOWS = [("channel", 300, 72, (60, 20, 90, 30)), ("channel", 131, 37, (1, 1, 128, 35)),
           ("periodic", 300, 72, (37, 13, 200, 41)), ("periodic", 97, 41, (5, 30, 64, 10))]


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("case,W,H,window", WINDOWS)
def test_window_single_step_and_rebase_bitwise(gpu, storage, fmt, case, W, H, window):
    """Single steps with a base window, and lbm_rebase_window at both parities, bit for bit against the oracle's
    16-bit mirror with the same row table and window (state words and the new per-cell base)."""
    sim = _wsim(storage, case, W, H, window)
    op = _op(sim)
    prev, par = sim.state()
    wb = sim.window_vel_host()
    for n in range(10):
        if n in (3, 6):
            sim.rebase()
            got, _ = sim.state()
            nb = O.rebase_window(W, H, prev, fmt, par, window + (wb,))
            wb = sim.window_vel_host()
            assert np.array_equal(nb, wb), f"rebase at parity {par}: base differs"
            assert np.array_equal(got, prev), f"rebase at parity {par}: {np.count_nonzero(got != prev)} words differ"
        sim.step(1)
        got, gpar = sim.state()
        ref = prev.copy()
        O.aa_step_h16(op, sim.mask_host, ref, fmt, sim.inlet_u_host, par, base=sim.row_base_host, window=window + (wb,))
        sel = O.observable(op, sim.mask_host, par ^ 1, ghost=True)
        assert np.array_equal(got[sel], ref[sel]), f"step {n}: {np.count_nonzero(got[sel] != ref[sel])} differ"
        prev, par = got, gpar


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("case,W,H,window", WINDOWS + [("channel", 301, 90, (2, 40, 297, 49))])
@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_window_pair_equals_single(gpu, storage, fmt, case, W, H, window, arith):
    """The two-step kernel with a base window (tiles that meet it take the per-cell form) bit-identical to single
    steps, every tile height, with the window re-derived between launches."""
    for chunk in (8, 12, 16, 24):
        a = _wsim(storage, case, W, H, window, rebase_every=4, arith=arith)
        b = _wsim(storage, case, W, H, window, rebase_every=4, arith=arith)
        sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
        for launch in range(3):
            a.step(4)
            b.step(4, k=2, cfg=(0, chunk))
            assert int(((a.f != b.f) & sel).sum()) == 0, f"chunk {chunk} launch {launch}"
            assert torch.equal(a._wvel[0], b._wvel[0])


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("window", [None, (60, 20, 90, 30)])
def test_rowbase_fused_default_composed(gpu, storage, arith, window):
    """lbm_step_fused_h16(_n) without a configuration where one tile launch does not cover the k steps (k = 3 and 4
    from an even state: a two-step launch, then single steps in place; odd states and k = 1: a copy, then single
    steps) equal single steps bit for bit, the last step's macros included."""
    a = _wsim(storage, window=window, arith=arith)
    b = _wsim(storage, window=window, arith=arith)
    op = _op(a)
    upd = torch.from_numpy(O.observable(op, a.mask_host, 0)[0]).to(a.device)

    def same(what):
        sel = torch.from_numpy(O.observable(op, a.mask_host, a.parity, ghost=True)).to(a.device)
        assert a.parity == b.parity and a.steps == b.steps
        assert int(((a.f != b.f) & sel).sum()) == 0, what

    for k, macros in ((4, False), (3, True), (3, False), (4, True)):
        a.step(k, macros=macros)
        b.step(k, k=k, macros=macros)
        same(f"k={k}")
        if macros:
            for x, y in ((a.rho, b.rho), (a.ux, b.ux), (a.uy, b.uy)):
                assert torch.equal(x[upd], y[upd]), f"k={k} macros"
    for start in (0, 1):  # k = 1 (a copy and one step), then two launches of three steps from one C call
        if a.parity != start:
            a.step(1)
            b.step(1)
        a.step(1, macros=True)
        b._fused(1, True)
        same(f"k=1 from parity {start}")
        for x, y in ((a.rho, b.rho), (a.ux, b.ux), (a.uy, b.uy)):
            assert torch.equal(x[upd], y[upd]), "k=1 macros"
        a.step(6)
        b.step(6, k=3)
        same(f"two launches of k=3 from parity {start ^ 1}")


def test_window_conversion_macros_and_checkpoint(gpu, tmp_path):
    from lbm.sim import Simulation
    sim = _wsim("f16", rebase_every=4)
    op = _op(sim)
    sim.step(7)  # one rebase behind it, odd parity
    h, par = sim.state()
    win = sim.window + (sim.window_vel_host(),)
    f = O.populations(h, O.F16, base=sim.row_base_host, parity=par, window=win)
    assert np.allclose(sim.populations().cpu().numpy(), f, rtol=0, atol=1e-7)
    rho, ux, uy = (t.cpu().numpy() for t in sim.macro())
    r, a, b = O.macro(op, sim.mask_host, sim.populations().cpu().numpy(), sim.inlet_u_host, par)
    upd = O.observable(op, sim.mask_host, 1)[0]
    assert np.array_equal(a[upd], ux[upd]) and np.array_equal(b[upd], uy[upd]) and np.array_equal(r[upd], rho[upd])
    sim.save(tmp_path / "w.npz")
    back = Simulation.load(tmp_path / "w.npz")
    assert back.window == sim.window and back.rebase_every == 4 and back._since_rebase == sim._since_rebase
    assert torch.equal(back.f, sim.f) and torch.equal(back._wvel[0], sim._wvel[0])
    sim.step(1)
    back.step(1)
    sim.step(8, k=2, cfg=(0, 12))
    back.step(8, k=2, cfg=(0, 12))
    assert torch.equal(back.f, sim.f) and torch.equal(back._wvel[0], sim._wvel[0])


def test_window_accuracy_scaled_cylinder(gpu):
    """The cylinder flow at 1024 x 256 (the 4096 x 1024 geometry scaled), 1000 steps: FP16 storage relative to the
    inlet profile with a base window over the near wake (re-derived every 200 steps) against FP64 is no worse than
    the row base alone -- printed for both."""
    from lbm.sim import Simulation
    case = configs.cylinder_scaled(1024, 256, steps=1000)
    steps = 1000
    p = _op(Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs()))
    f64 = case.f0.astype(np.float64).copy()
    par = O.aa_steps(p, case.mask, f64, np.asarray(case.inlet_u, np.float64), 0, steps)
    R, A, B = O.macro(p, case.mask, f64, np.asarray(case.inlet_u, np.float64), par)
    upd = O.observable(p, case.mask, 1)[0]
    errs = {}
    for win in (None, configs.wake_window(case)):
        sim = Simulation(case.W, case.H, case.tau_inv, storage="f16", row_base="inlet", window=win,
                         **case.sim_kwargs())
        sim.step(steps, k=2)
        r, a, b = (t.cpu().numpy().astype(np.float64) for t in sim.macro())
        errs[win] = np.sqrt((((a - A) ** 2 + (b - B) ** 2)[upd]).sum() / ((A ** 2 + B ** 2)[upd]).sum())
    print(f"\nL2(u) vs FP64 at {steps} steps: {errs}")
    assert errs[configs.wake_window(case)] <= errs[None] * 1.05, errs


def test_window_pair_cylinder_large(gpu):
    """The 4096 x 1024 cylinder's 10 000 steps with FP16 storage relative to the inlet profile, the wake's base window
    re-derived every 200 steps and the fast arithmetic (the bench's headline form): the 12-row two-step kernel
    bit-identical to single steps, window velocities included."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    kw = dict(storage="f16", row_base="inlet", arith="fast", window=configs.wake_window(case), rebase_every=200,
              **case.sim_kwargs())
    a = Simulation(case.W, case.H, case.tau_inv, **kw)
    b = Simulation(case.W, case.H, case.tau_inv, **kw)
    sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
    for chunk in range(5):
        a.step(2000)
        b.step(2000, k=2, cfg=(0, 12))
        n = int(((a.f != b.f) & sel).sum())
        assert n == 0, f"after {2000 * (chunk + 1)} steps: {n} differ"
        assert torch.equal(a._wvel[0], b._wvel[0])
    assert np.isfinite(b.diagnostics()).all()


def test_window_wake_keyword(gpu):
    """Simulation(window="wake") takes the window of the mask's bodies (geometry.obstacle_wake_window); the
    two-step kernel then equals single steps."""
    from lbm.sim import Simulation
    W, H = 300, 72
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 6))
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.12, mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, 6), storage="f16",
              row_base="inlet", window="wake", rebase_every=4, arith="fast")
    a = Simulation(W, H, 1.75, **kw)
    b = Simulation(W, H, 1.75, **kw)
    assert a.window == G.obstacle_wake_window(a.mask_host) == (75, 24, 24, 24)
    sel = torch.from_numpy(O.observable(_op(a), a.mask_host, 0, ghost=True)).to(a.device)
    a.step(8)
    b.step(8, k=2, cfg=(0, 12))
    assert int(((a.f != b.f) & sel).sum()) == 0
    with pytest.raises(ValueError, match="no body"):
        Simulation(W, H, 1.75, storage="f16", row_base="inlet", window="wake")


@pytest.mark.parametrize("model,sums", [("bgk", "plain"), ("mrt", "neumaier")])
def test_window_pair_other_models(gpu, model, sums):
    """The window forms of the BGK and compensated-sum kernels: single steps bitwise = the mirror, the two-step
    kernel bitwise = single steps."""
    from lbm.sim import Simulation
    W, H, win = 301, 90, (2, 40, 297, 49)
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 6))
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.0 if model == "bgk" else 0.12, mask=mask, inlet_u=prof, model=model, sums=sums,
              f0=G.channel_initial_state(W, H, prof, 1e-3, 6), storage="f16", row_base="inlet", window=win,
              rebase_every=4)
    a = Simulation(W, H, 1.75, **kw)
    b = Simulation(W, H, 1.75, **kw)
    op = _op(a)
    prev, par = a.state()
    wv = a.window_vel_host()
    for _ in range(4):
        a.step(1)
        O.aa_step_h16(op, a.mask_host, prev, O.F16, a.inlet_u_host, par, base=a.row_base_host, window=win + (wv,))
        par ^= 1
        if a._since_rebase == 0:  # re-derived after this step
            wv = O.rebase_window(W, H, prev, O.F16, par, win + (wv,))
        sel = O.observable(op, a.mask_host, par, ghost=True)
        assert np.array_equal(a.state()[0][sel], prev[sel])
    b.step(4, k=2, cfg=(0, 12))
    sel = torch.from_numpy(O.observable(op, a.mask_host, 0, ghost=True)).to(a.device)
    assert int(((a.f != b.f) & sel).sum()) == 0
