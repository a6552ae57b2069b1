"""GPU: compensated (Neumaier) sums of rho - 1 and j (lbm_params.sums = 1) in every kernel that advances the
lattice -- the strip and tile kernels of lbm_step_fused (k = 1..8, two- and four-step tiles, FP32 and 16-bit
storage), the persistent ensemble kernels (k = 1, 2, 4) and the two-buffer pull step -- bit for bit against
single compensated AA steps (which the single-step tests hold bit for bit to the oracle), on flows strong
enough that the compensated sums differ from the plain ones."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from lbm import geometry as G

pytestmark = pytest.mark.gpu


def _strong_periodic(W, H, seed, storage="f32", sums="neumaier", tau_inv=1.7):
    """u ~ 0.2 and |rho - 1| ~ 0.05: partial sums above 2^-5, where plain FP32 sums round."""
    from lbm.sim import Simulation
    rng = np.random.default_rng(seed)
    ux, uy = G.random_phase_velocity(W, H, k0=4, u_rms=0.2, seed=seed)
    f0 = G.feq(1.0 + 0.05 * rng.standard_normal((H, W)), ux, uy)
    mask = G.periodic_mask(W, H)
    if W > 16 and H > 16:
        mask[H // 3:H // 3 + 5, W // 2:W // 2 + 6] = 1
    return Simulation(W, H, tau_inv, cs=0.1, bc="periodic", mask=mask, f0=f0, storage=storage, sums=sums)


def _channel(W, H, seed, storage="f32", sums="neumaier", tau_inv=1.75):
    from lbm.sim import Simulation
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, max(2, H // 10)))
    prof = G.parabolic_profile(H, 0.08)
    f0 = G.channel_initial_state(W, H, prof, 1e-2, seed)
    return Simulation(W, H, tau_inv, cs=0.12, bc="channel", mask=mask, inlet_u=prof, f0=f0, storage=storage,
                      sums=sums)


def _sel(sim):
    p = sim.params
    op = O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc)
    return [torch.from_numpy(O.observable(op, sim.mask_host, par, ghost=True)).to(sim.device) for par in (0, 1)]


def _bits(t):
    return t.view(torch.int16 if t.dtype in (torch.int16, torch.float16, torch.bfloat16) else torch.int32)


def _assert_equal(a, b, sel, what):
    assert a.parity == b.parity, what
    n = int(((_bits(a.f) != _bits(b.f)) & sel[a.parity]).sum())
    assert n == 0, f"{what}: {n} populations differ"


def _make(periodic, W, H, seed, storage="f32", sums="neumaier"):
    return (_strong_periodic if periodic else _channel)(W, H, seed, storage=storage, sums=sums)


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("periodic", [True, False])
def test_strip_kernel_sums_equals_single(gpu, k, periodic):
    W, H = (300, 70) if periodic else (333, 97)
    a, b = _make(periodic, W, H, 5), _make(periodic, W, H, 5)
    sel = _sel(a)
    for start_parity in (0, 1):
        if start_parity:
            a.step(1)
            b.step(1)
        for launch in range(3):
            a.step(k)
            b.step(k, k=k, cfg=(128, 24))
            _assert_equal(a, b, sel, f"k={k} launch {launch} start parity {start_parity}")


@pytest.mark.parametrize("storage,chunk", [("f32", 8), ("f32", 24), ("f16", 12), ("bf16", 16)])
@pytest.mark.parametrize("periodic", [True, False])
def test_pair_kernel_sums_equals_two_single_steps(gpu, storage, chunk, periodic):
    for W, H in ([(300, 70), (66, 44)] if periodic else [(333, 97), (134, 16)]):
        a, b = _make(periodic, W, H, 7, storage), _make(periodic, W, H, 7, storage)
        sel = _sel(a)
        for launch in range(3):
            a.step(2)
            b.step(2, k=2, cfg=(0, chunk))
            _assert_equal(a, b, sel, f"{storage} {W}x{H} chunk {chunk} launch {launch}")


@pytest.mark.parametrize("chunk", [8, 12, 14, 16, 20, 24, 32])
@pytest.mark.parametrize("periodic", [True, False])
def test_quad_kernel_sums_equals_four_single_steps(gpu, chunk, periodic):
    for W, H in ([(300, 70), (40, 12)] if periodic else [(333, 97), (70, 9)]):
        a, b = _make(periodic, W, H, 9), _make(periodic, W, H, 9)
        sel = _sel(a)
        for launch in range(3):
            a.step(4)
            b.step(4, k=4, cfg=(0, chunk))
            _assert_equal(a, b, sel, f"{W}x{H} chunk {chunk} launch {launch}")


@pytest.mark.parametrize("k,cfg", [(2, (0, 12)), (4, (0, 24)), (3, (128, 24))])
def test_blocked_sums_mode_is_live(gpu, k, cfg):
    """The same blocked launches with plain sums end in a different state: the compensated instantiation ran."""
    W, H = 256, 64
    comp, plain = _strong_periodic(W, H, 3), _strong_periodic(W, H, 3, sums="plain")
    comp.step(k, k=k, cfg=cfg)
    plain.step(k, k=k, cfg=cfg)
    assert not torch.equal(_bits(comp.f), _bits(plain.f))


def test_pull_step_sums_bitwise(gpu):
    """The two-buffer pull step with compensated sums against the oracle's pull step in the same mode."""
    from lbm import _lib
    sim = _strong_periodic(96, 40, 13)
    p = sim.params
    op = O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp, sums=p.sums)
    assert p.sums == 1
    L = _lib.load()
    f_now = sim.f.clone()
    f_out = f_now.clone()
    ux = torch.zeros((sim.H, sim.W), device=sim.device)
    uy, rho = torch.zeros_like(ux), torch.zeros_like(ux)
    ref = f_now.cpu().numpy()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in range(6):
        ptrs = [ctypes.c_void_p(t.data_ptr()) for t in (sim.mask, f_now, f_out, ux, uy, rho, sim.inlet_u)]
        rc = L.lbm_step_pull(ctypes.byref(p), *ptrs, st)
        _lib.check(rc, "lbm_step_pull")
        ref_out = ref.copy()
        O.pull_step(op, sim.mask_host, ref, ref_out, sim.inlet_u_host)
        sel = O.observable(op, sim.mask_host, 1)
        got = f_out.cpu().numpy()
        assert np.array_equal(got[sel], ref_out[sel]), f"step {n}"
        f_now, f_out = f_out, f_now
        ref = got


SIZES = [(64, 32), (70, 33), (129, 5), (256, 130), (3, 3), (200, 64)]


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_ensemble_sums_bitwise(gpu, storage):
    """Persistent k = 1, 2 (and 4 for FP32) launches over compensated-sum members equal each member stepped
    alone with single compensated steps."""
    from lbm.ensemble import Ensemble
    a = [_strong_periodic(W, H, 20 + i, storage) for i, (W, H) in enumerate(SIZES)]
    b = [_strong_periodic(W, H, 20 + i, storage) for i, (W, H) in enumerate(SIZES)]
    sels = [_sel(s) for s in b]
    ks = [1, 2, 4] if storage == "f32" else [1, 2]
    ens = Ensemble(a, two_step=True, k=2, chunk=8 if storage == "f32" else 12)
    for k in ks:
        ens.k = k
        if k == 4:
            ens.chunk = 20
        ens.step(2 * k)
        for s in b:
            s.step(2 * k)
        torch.cuda.synchronize()
        for i, (x, y) in enumerate(zip(a, b)):
            _assert_equal(x, y, sels[i], f"{storage} k={k} member {i} ({x.W}x{x.H})")


def test_ensemble_sums_members_must_agree(gpu):
    from lbm import _lib
    from lbm.ensemble import Ensemble
    with pytest.raises(_lib.LbmError):
        Ensemble([_strong_periodic(64, 32, 1), _strong_periodic(64, 32, 2, sums="plain")])
