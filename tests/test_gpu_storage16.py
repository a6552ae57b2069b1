"""GPU: FP16/BF16 deviation storage -- single steps bit for bit against the oracle's 16-bit mirror, the
k-step kernel bit for bit against 16-bit single steps, conversions, checkpoints."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm import geometry as G

pytestmark = pytest.mark.gpu

FMTS = [("f16", O.F16), ("bf16", O.BF16)]


def _sim(storage, periodic=False, W=300, H=72, cs=0.12, seed=6, sums="plain"):
    """W = 300: the 128 x 2 tiles at x = 128..255 are interior (single-step interior form), with an obstacle."""
    from lbm.sim import Simulation
    if periodic:
        mask = G.periodic_mask(W, H)
        mask[20:25, 60:66] = 1
        mask[40:47, 180:186] = 1
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=seed)
        return Simulation(W, H, 1.7, cs=cs, bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy),
                          storage=storage, sums=sums)
    mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, 6))
    mask[20:26, 200:206] = 1
    prof = G.parabolic_profile(H, 0.05)
    return Simulation(W, H, 1.75, cs=cs, mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed),
                      storage=storage, sums=sums)


def _op(sim):
    p = sim.params
    return O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp, sums=p.sums)


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("sums", ["plain", "neumaier"])
def test_h16_single_step_bitwise(gpu, storage, fmt, periodic, sums):
    sim = _sim(storage, periodic, sums=sums)
    op = _op(sim)
    prev, par = sim.state()
    for n in range(10):
        sim.step(1)
        got, gpar = sim.state()
        ref = prev.copy()
        O.aa_step_h16(op, sim.mask_host, ref, fmt, sim.inlet_u_host, par)
        sel = O.observable(op, sim.mask_host, par ^ 1, ghost=True)
        assert gpar == par ^ 1
        assert np.array_equal(got[sel], ref[sel]), f"step {n}: {np.count_nonzero(got[sel] != ref[sel])} differ"
        prev, par = got, gpar


@pytest.mark.parametrize("storage,fmt", FMTS)
@pytest.mark.parametrize("k,chunk", [(2, 16), (4, 23), (8, 64), (3, 40)])
def test_h16_fused_equals_single(gpu, storage, fmt, k, chunk):
    a = _sim(storage, W=301, H=90)
    b = _sim(storage, W=301, H=90)
    op = _op(a)
    sels = [torch.from_numpy(O.observable(op, a.mask_host, q, ghost=True)).to(a.device) for q in (0, 1)]
    for launch in range(3):
        a.step(k)
        b.step(k, k=k, cfg=(128, chunk))
        assert a.parity == b.parity
        n = int(((a.f != b.f) & sels[a.parity]).sum())
        assert n == 0, f"k={k} launch {launch}: {n} differ"


@pytest.mark.parametrize("storage,fmt", FMTS)
def test_h16_conversion_and_macros(gpu, storage, fmt):
    sim = _sim(storage)
    f0 = G.channel_initial_state(sim.W, sim.H, G.parabolic_profile(sim.H, 0.05), 1e-3, 6).astype(np.float32)
    h = sim.state()[0]
    want = O.to_h16(f0 - O.W_F32[:, None, None], fmt)
    assert np.array_equal(h, want)
    sim.step(7)
    rho, ux, uy = (t.cpu().numpy() for t in sim.macro())
    op = _op(sim)
    h, par = sim.state()
    f = (O.from_h16(h, fmt) + O.W_F32[:, None, None]).astype(np.float32)
    assert np.array_equal(sim.populations().cpu().numpy(), f)
    r, a, b = O.macro(op, sim.mask_host, f, sim.inlet_u_host, par)
    upd = O.observable(op, sim.mask_host, 1)[0]
    assert np.array_equal(r[upd], rho[upd]) and np.array_equal(a[upd], ux[upd]) and np.array_equal(b[upd], uy[upd])


def test_checkpoint_roundtrip(gpu, tmp_path):
    from lbm.sim import Simulation
    for storage in ("f32", "f16"):
        sim = _sim(storage)
        sim.step(5)
        sim.save(tmp_path / f"c_{storage}.npz")
        back = Simulation.load(tmp_path / f"c_{storage}.npz")
        assert back.parity == sim.parity and back.steps == 5 and torch.equal(back.f, sim.f)
        sim.step(3)
        back.step(3)
        assert torch.equal(back.f, sim.f)
