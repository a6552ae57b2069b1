"""GPU parity of the single-step kernels against the oracle (FP32 mirror bit for bit, FP64 within the
stated tolerance).  Everything goes through liblbm's C-ABI via lbm.Simulation / ctypes."""
import ctypes

import numpy as np
import pytest

import oracle as O
from lbm import configs
from lbm import geometry as G

pytestmark = pytest.mark.gpu

REL_TOL_FP64 = 1e-6  # FP32 single step vs FP64 oracle, per population, per step


def _oparams(case_or_sim, **kw):
    s = case_or_sim
    p = s.params
    return O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp, s_e=p.s_e, s_eps=p.s_eps,
                    s_q=p.s_q, sums=p.sums)


def _sim(case, **kw):
    from lbm.sim import Simulation
    return Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs(), **kw)


def _small_channel(W=96, H=40, model="mrt", cs=0.1, seed=4, **kw):
    from lbm.sim import Simulation
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 5))
    prof = G.parabolic_profile(H, 0.05)
    f0 = G.channel_initial_state(W, H, prof, noise=1e-3, seed=seed)
    return Simulation(W, H, 1.8, cs=cs, model=model, bc="channel", mask=mask, inlet_u=prof, f0=f0, **kw)


def _small_periodic(W=80, H=48, cs=0.12, model="mrt", **kw):
    from lbm.sim import Simulation
    mask = G.periodic_mask(W, H)
    mask[20:26, 30:37] = 1
    ux, uy = G.random_phase_velocity(W, H, k0=4, u_rms=0.04, seed=5)
    return Simulation(W, H, 1.7, cs=cs, model=model, bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy),
                      **kw)


def _check_steps_bitwise(gpu, sim, nsteps, macros=False):
    """Every step: the GPU's output equals the FP32 mirror applied to the GPU's previous state."""
    p = _oparams(sim)
    mask = sim.mask_host
    prof = sim.inlet_u_host
    f_prev, par = sim.state()
    for n in range(nsteps):
        sim.step(1, macros=macros)
        f_gpu, par_gpu = sim.state()
        ref = f_prev.copy()
        mac = O.aa_step(p, mask, ref, prof, par, macros=macros)
        sel = O.observable(p, mask, par ^ 1, ghost=True)
        assert par_gpu == par ^ 1
        bad = np.count_nonzero(f_gpu[sel] != ref[sel])
        assert bad == 0, f"step {n} parity {par}: {bad} populations differ, max {np.abs(f_gpu - ref)[sel].max()}"
        if macros:
            upd = sel[0] | O.observable(p, mask, 1)[0]
            for a, b in zip(mac, (sim.rho, sim.ux, sim.uy)):
                assert np.array_equal(a[upd], b.cpu().numpy()[upd]), f"macros step {n}"
        f_prev, par = f_gpu, par_gpu


@pytest.mark.parametrize("model,cs", [("mrt", 0.1), ("mrt", 0.0), ("bgk", 0.0)])
def test_aa_step_bitwise_channel(gpu, model, cs):
    _check_steps_bitwise(gpu, _small_channel(model=model, cs=cs), 12)


def test_aa_step_macros_bitwise(gpu):
    _check_steps_bitwise(gpu, _small_channel(), 6, macros=True)


@pytest.mark.parametrize("model", ["mrt", "bgk"])
def test_aa_step_bitwise_periodic(gpu, model):
    _check_steps_bitwise(gpu, _small_periodic(model=model, cs=0.12 if model == "mrt" else 0.0), 12)


@pytest.mark.parametrize("periodic", [False, True])
def test_aa_step_bitwise_interior_tiles(gpu, periodic):
    """Lattices wide enough for interior 128 x 2 tiles (x = 128..255 here, taking the single-step kernel's
    interior form), with obstacles inside and across them: every step equals the FP32 mirror."""
    W, H = 400, 40
    if periodic:
        sim = _small_periodic(W=W, H=H)
    else:
        sim = _small_channel(W=W, H=H)
    sim.mask[10:16, 150:160] = 1  # inside an interior tile
    sim.mask[22:27, 250:262] = 1  # across the x = 256 tile edge
    sim.mask_host[10:16, 150:160] = 1
    sim.mask_host[22:27, 250:262] = 1
    _check_steps_bitwise(gpu, sim, 10)


def test_aa_step_odd_sizes(gpu):
    """Lattices that do not fill the 64 x 4 tiles (ragged right and bottom edges)."""
    from lbm.sim import Simulation
    for W, H in ((3, 3), (65, 5), (131, 7), (70, 129)):
        mask = G.channel_mask(W, H)
        prof = G.parabolic_profile(H, 0.04)
        f0 = G.channel_initial_state(W, H, prof, noise=1e-3, seed=W)
        sim = Simulation(W, H, 1.6, cs=0.1, mask=mask, inlet_u=prof, f0=f0)
        _check_steps_bitwise(gpu, sim, 4)


def test_cylinder_small_1000_steps_vs_fp64(gpu):
    """512 x 128 cylinder, 1000 single AA steps: at every step the GPU's FP32 output is within 1e-6
    relative per population of the FP64 oracle advanced one step from the GPU's previous state, and equal
    bit for bit to the FP32 mirror.  The free-running FP32 vs FP64 drift is reported, not gated."""
    case = configs.cylinder_small()
    sim = _sim(case)
    p = _oparams(sim)
    mask, prof32 = sim.mask_host, sim.inlet_u_host
    prof64 = prof32.astype(np.float64)
    f_prev, par = sim.state()
    free64 = f_prev.astype(np.float64)
    worst = 0.0
    for n in range(case.steps):
        sim.step(1)
        f_gpu, par_gpu = sim.state()
        sel = O.observable(p, mask, par ^ 1)
        r64 = f_prev.astype(np.float64)
        O.aa_step(p, mask, r64, prof64, par)
        rel = np.abs(f_gpu[sel] - r64[sel]) / np.abs(r64[sel])
        worst = max(worst, float(rel.max()))
        assert rel.max() <= REL_TOL_FP64, f"step {n}: {rel.max():.3e}"
        if n % 50 == 0 or n == case.steps - 1:
            r32 = f_prev.copy()
            O.aa_step(p, mask, r32, prof32, par)
            assert np.array_equal(f_gpu[sel], r32[sel]), f"step {n}: FP32 mirror differs"
        O.aa_step(p, mask, free64, prof64, par)
        f_prev, par = f_gpu, par_gpu
    sel = O.observable(p, mask, par)
    drift = np.abs(f_prev[sel] - free64[sel]) / np.abs(free64[sel])
    print(f"\nper-step worst rel err {worst:.3e}; free-running 1000-step drift max {drift.max():.3e}")


def test_pull_step_bitwise(gpu):
    """Reference-compatible two-buffer step (lbm_core4.cu semantics) vs the FP32 mirror, with macros."""
    import torch
    from lbm import _lib
    sim = _small_channel()
    p = _oparams(sim)
    L = _lib.load()
    f_now = sim.f.clone()
    f_out = f_now.clone()
    ux = torch.zeros((sim.H, sim.W), device=sim.device)
    uy = torch.zeros_like(ux)
    rho = torch.zeros_like(ux)
    ref = f_now.cpu().numpy()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in range(8):
        rc = L.lbm_step_pull(ctypes.byref(sim.params), ctypes.c_void_p(sim.mask.data_ptr()),
                             ctypes.c_void_p(f_now.data_ptr()), ctypes.c_void_p(f_out.data_ptr()),
                             ctypes.c_void_p(ux.data_ptr()), ctypes.c_void_p(uy.data_ptr()),
                             ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(sim.inlet_u.data_ptr()), st)
        _lib.check(rc, "lbm_step_pull")
        ref_out = ref.copy()
        mac = O.pull_step(p, sim.mask_host, ref, ref_out, sim.inlet_u_host)
        sel = O.observable(p, sim.mask_host, 1)
        got = f_out.cpu().numpy()
        assert np.array_equal(got[sel], ref_out[sel]), f"step {n}"
        for a, b in zip(mac, (rho, ux, uy)):
            assert np.array_equal(a[sel[0]], b.cpu().numpy()[sel[0]])
        f_now, f_out = f_out, f_now
        ref = got


@pytest.mark.parametrize("parity", [0, 1])
def test_macro_bitwise(gpu, parity):
    sim = _small_channel()
    sim.step(5 + parity)
    rho, ux, uy = (t.cpu().numpy() for t in sim.macro())
    f, par = sim.state()
    assert par == (5 + parity) & 1
    p = _oparams(sim)
    r, a, b = O.macro(p, sim.mask_host, f, sim.inlet_u_host, par)
    upd = O.observable(p, sim.mask_host, 1)[0]
    assert np.array_equal(r[upd], rho[upd]) and np.array_equal(a[upd], ux[upd]) and np.array_equal(b[upd], uy[upd])


def test_visualize_matches_numpy(gpu):
    sim = _small_channel()
    sim.step(20)
    img = sim.visualize(50.0).cpu().numpy()
    ux, uy = sim.ux.cpu().numpy(), sim.uy.cpu().numpy()
    H, W = ux.shape
    vort = ((uy[1:-1, 2:] - uy[1:-1, :-2]) - (ux[2:, 1:-1] - ux[:-2, 1:-1])) * np.float32(50.0)
    one = np.float32(1)
    r = np.minimum(np.maximum(one + vort, 0), 1)
    g = np.minimum(np.maximum(one - np.abs(vort), 0), 1)
    b = np.minimum(np.maximum(one - vort, 0), 1)
    want = np.stack([(r * np.float32(255)).astype(np.uint8), (g * np.float32(255)).astype(np.uint8),
                     (b * np.float32(255)).astype(np.uint8)], -1)
    solid = sim.mask_host[1:-1, 1:-1].astype(bool)
    want[solid] = 0
    assert np.array_equal(img[1:-1, 1:-1, :3], want)
    assert np.all(img[1:-1, 1:-1, 3] == 255)


@pytest.mark.parametrize("parity", [0, 1])
def test_diagnostics_fp64(gpu, parity):
    sim = _small_channel(W=200, H=64)
    sim.step(30 + parity)
    d = sim.diagnostics()
    f, par = sim.state()
    ref = O.diagnostics(_oparams(sim), sim.mask_host, f, sim.inlet_u_host, par)
    assert abs(d[0] - ref[0]) <= 1e-12 * abs(ref[0])
    assert np.all(np.abs(d[1:] - ref[1:]) <= 1e-12 * (1 + np.abs(ref[1:])))
    d2 = sim.diagnostics()
    assert np.array_equal(d, d2)  # deterministic reduction


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_aa_step_bitwise_clamps_engaged(gpu, storage):
    """Densities from -0.3 to 2.1 and speeds up to 0.6: the velocity clamp and both ends of the density clamp
    (rho -> 0, where 1/rho_c = +inf, and rho -> 1.5) engage on many cells; every step still equals the FP32
    (or 16-bit) mirror bit for bit."""
    from lbm.sim import Simulation
    W, H = 96, 40
    rng = np.random.default_rng(12)
    rho = rng.uniform(-0.3, 2.1, (H, W))
    ux = rng.uniform(-0.6, 0.6, (H, W))
    uy = rng.uniform(-0.6, 0.6, (H, W))
    sim = Simulation(W, H, 1.7, cs=0.15, model="mrt", bc="periodic", mask=G.periodic_mask(W, H),
                     f0=G.feq(rho, ux, uy), storage=storage)
    p = _oparams(sim)
    zero = np.zeros(H, np.float32)
    if storage == "f32":
        f_prev, par = sim.state()
        for n in range(4):
            sim.step(1)
            ref = f_prev.copy()
            O.aa_step(p, sim.mask_host, ref, zero, par)
            got, par2 = sim.state()
            sel = O.observable(p, sim.mask_host, par2)
            both = np.isnan(got[sel]) & np.isnan(ref[sel])
            assert np.array_equal(got[sel][~both], ref[sel][~both]), f"step {n}"
            f_prev, par = got, par2
    else:
        h_prev = sim.f.cpu().numpy().view(np.uint16).copy()
        par = sim.parity
        for n in range(4):
            sim.step(1)
            ref = h_prev.copy()
            O.aa_step_h16(p, sim.mask_host, ref, O.F16, zero, par)
            got = sim.f.cpu().numpy().view(np.uint16)
            sel = O.observable(p, sim.mask_host, par ^ 1)
            assert np.array_equal(got[sel], ref[sel]), f"step {n}"
            h_prev, par = got.copy(), par ^ 1


def test_health_check_detects_injected_failures(gpu):
    """Failure detection, forced from the host: a NaN, a negative density and a mass drift injected into a
    healthy state are each reported; the healthy state passes."""
    import torch
    from lbm.sim import LbmHealthError
    sim = _small_periodic()
    mass0 = sim.diagnostics()[0]
    sim.step(10)
    ok = sim.check_health(mass0, rel_tol=1e-6)
    assert ok["mass_drift"] < 1e-6 and ok["rho_min"] > 0.5
    for inject, what in (("nan", "non-finite"), ("neg", "negative"), ("scale", "drift")):
        s = _small_periodic()
        s.step(4)
        with torch.no_grad():
            if inject == "nan":
                s.f[3, 10, 11] = float("nan")
            elif inject == "neg":
                s.f[:, 10, 11] = -0.5
            else:
                s.f.mul_(1.01)
        with pytest.raises(LbmHealthError, match=what):
            s.check_health(mass0, rel_tol=1e-3)


@pytest.mark.parametrize("periodic", [False, True])
def test_aa_step_compensated_sums_bitwise(gpu, periodic):
    """sums = 1 (Neumaier sums of rho - 1 and j): every step equals the FP32 mirror in the same mode bit for bit,
    and differs from the plain-sum step (the mode is live)."""
    sim = (_small_periodic if periodic else _small_channel)(sums="neumaier")
    assert sim.params.sums == 1
    _check_steps_bitwise(gpu, sim, 10, macros=True)
    if periodic:  # a strong flow, where the plain sums round: the compensated steps differ and still match
        from lbm.sim import Simulation
        W, H = 64, 32
        rng = np.random.default_rng(3)
        ux, uy = G.random_phase_velocity(W, H, 4, 0.25, seed=3)
        f0 = G.feq(1.0 + 0.08 * rng.standard_normal((H, W)), ux, uy)
        kw = dict(cs=0.1, bc="periodic", mask=G.periodic_mask(W, H), f0=f0)
        plain, comp = Simulation(W, H, 1.7, **kw), Simulation(W, H, 1.7, sums="neumaier", **kw)
        _check_steps_bitwise(gpu, comp, 4)
        plain.step(4)
        assert not np.array_equal(plain.f.cpu().numpy(), comp.f.cpu().numpy())
