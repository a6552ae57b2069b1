"""GPU: the temporally blocked kernel equals k single AA steps bit for bit (observable state)."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm import configs
from lbm import geometry as G

pytestmark = pytest.mark.gpu


def _pair(W, H, periodic, model, seed=3):
    from lbm.sim import Simulation
    if periodic:
        mask = G.periodic_mask(W, H)
        mask[H // 3:H // 3 + 5, W // 2:W // 2 + 6] = 1
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=seed)
        kw = dict(bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy))
    else:
        mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, max(2, H // 10)))
        prof = G.parabolic_profile(H, 0.05)
        kw = dict(bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed))
    cs = 0.12 if model == "mrt" else 0.0
    return (Simulation(W, H, 1.75, cs=cs, model=model, **kw), Simulation(W, H, 1.75, cs=cs, model=model, **kw))


def _sel(sim):
    p = sim.params
    op = O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc)
    return [torch.from_numpy(O.observable(op, sim.mask_host, par, ghost=True)).to(sim.device) for par in (0, 1)]


def _assert_equal(a, b, sel, what):
    diff = (a.f != b.f) & sel[a.parity]
    n = int(diff.sum())
    assert a.parity == b.parity
    assert n == 0, f"{what}: {n} populations differ"


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("periodic,model", [(False, "mrt"), (True, "mrt"), (False, "bgk")])
def test_fused_equals_single_small(gpu, k, periodic, model):
    W, H = (333, 97) if not periodic else (300, 70)
    single, fusedsim = _pair(W, H, periodic, model)
    sel = _sel(single)
    for start_parity in (0, 1):
        if start_parity == 1:
            single.step(1)
            fusedsim.step(1)
        for launch in range(3):
            single.step(k)
            fusedsim.step(k, k=k)
            _assert_equal(single, fusedsim, sel, f"k={k} launch {launch} start parity {start_parity}")
        single.step(4 * k)
        fusedsim.step(4 * k, k=k)  # four launches from one C call (lbm_step_fused_n)
        _assert_equal(single, fusedsim, sel, f"k={k} four launches of one call, start parity {start_parity}")


@pytest.mark.parametrize("k", [1, 3, 6, 8])
def test_fused_default_fast_composed(gpu, k):
    """FP32 populations with the fast arithmetic and no configuration, k steps that no single tile launch covers:
    a four- or two-step launch (even states) or a copy (odd states, k = 1), then single steps in place, equal to
    single fast steps bit for bit, from both parities and over two launches of one C call."""
    from lbm.sim import Simulation
    W, H = 301, 90
    mask = G.channel_mask(W, H, cylinder=(70, 45, 9))
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.12, bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, 4),
              arith="fast")
    single, fusedsim = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    sel = _sel(single)
    for start in (0, 1):
        if single.parity != start:
            single.step(1)
            fusedsim.step(1)
        single.step(k)
        fusedsim._fused(k)
        _assert_equal(single, fusedsim, sel, f"k={k} from parity {start}")
        single.step(2 * k)
        fusedsim._fused_n(k, 2)
        _assert_equal(single, fusedsim, sel, f"two launches of k={k}")


@pytest.mark.parametrize("k,cfg", [(3, (64, 24)), (8, (64, 7)), (8, (128, 33)), (2, (64, 256)), (5, (128, 16)),
                                   (2, (128, 32)), (4, (128, 32)), (4, (64, 40))])
def test_fused_explicit_configs(gpu, k, cfg):
    single, fusedsim = _pair(301, 90, False, "mrt")
    sel = _sel(single)
    for launch in range(2):
        single.step(k)
        fusedsim.step(k, k=k, cfg=cfg)
        _assert_equal(single, fusedsim, sel, f"k={k} cfg={cfg} launch {launch}")


def test_autotuner_picks_a_valid_config(gpu, tmp_path, monkeypatch):
    from lbm import autotune
    monkeypatch.setenv("LBM_TUNE_CACHE", str(tmp_path / "tune.json"))
    single, fusedsim = _pair(512, 256, True, "mrt")
    before = fusedsim.f.clone()
    best, table = autotune.tune_fused(fusedsim, ks=(2, 4))
    assert torch.equal(before, fusedsim.f)  # tuning works on scratch copies
    k, strip, chunk, us = best
    assert k in (2, 4) and strip in (0, 64, 128) and us > 0
    sel = _sel(single)
    single.step(2 * k)
    fusedsim.step(2 * k, k=k, cfg=(strip, chunk))
    _assert_equal(single, fusedsim, sel, "tuned config")
    best2, _ = autotune.tune_fused(fusedsim, ks=(2, 4))
    assert best2 == best  # served from the cache


def test_autotuner_candidates_by_parity(gpu):
    """Even states: tile kernels only (k = 2; k = 4 for FP32), none for other k; odd states: the library's own
    choice (None, None) and the strip shapes."""
    from lbm import autotune
    _, sim = _pair(512, 256, True, "mrt")
    assert sim.parity == 0
    assert {s for s, _ in autotune.candidates(sim, 2)} == {0}
    assert {s for s, _ in autotune.candidates(sim, 4)} == {0}
    assert autotune.candidates(sim, 6) == []
    sim.step(1)
    odd = autotune.candidates(sim, 4)
    assert odd[0] == (None, None) and len(odd) > 1 and all(s in (64, 128) for s, _ in odd[1:])
    best, table = autotune.tune_fused(sim, ks=(2, 4), use_cache=False, reps=2)
    assert best[1] in (None, 64, 128) and len(table) == 2
    single = _pair(512, 256, True, "mrt")[0]
    single.step(1)
    sel = _sel(single)
    single.step(2 * best[0])
    sim.step(2 * best[0], k=best[0], cfg=best[1:3])  # (None, None) runs without a configuration
    _assert_equal(single, sim, sel, f"tuned {best}")


def test_fused_macros_match_single(gpu):
    single, fusedsim = _pair(200, 64, False, "mrt")
    single.step(3)
    single.step(1, macros=True)
    fusedsim.step(4, k=4, macros=True)
    upd = _sel(single)[1][0]
    for a, b in ((single.rho, fusedsim.rho), (single.ux, fusedsim.ux), (single.uy, fusedsim.uy)):
        assert torch.equal(a[upd], b[upd])


@pytest.mark.parametrize("k", [2, 4, 8])
def test_fused_equals_single_cylinder_large(gpu, k):
    """4096 x 1024 cylinder, 10000 steps: bitwise equal to single steps at every launch boundary."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    a = Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs())
    b = Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs())
    sel = _sel(a)
    for launch in range(case.steps // k):
        a.step(k)
        b.step(k, k=k)
        if launch % 25 == 0 or launch == case.steps // k - 1:
            _assert_equal(a, b, sel, f"launch {launch}")
    d = a.diagnostics()
    assert np.isfinite(d).all()


@pytest.mark.parametrize("chunk", [8, 12, 16, 24])
@pytest.mark.parametrize("periodic,model", [(False, "mrt"), (True, "mrt"), (False, "bgk"), (True, "bgk")])
def test_pair_kernel_equals_two_single_steps(gpu, chunk, periodic, model):
    """The two-step tile kernel (strip 0) equals two single steps bit for bit, on ragged sizes and on
    periodic lattices narrower than one 130-column tile region."""
    sizes = [(333, 97), (128, 16), (70, 9)] if not periodic else [(300, 70), (100, 30), (40, 12)]
    for W, H in sizes:
        single, fusedsim = _pair(W, H, periodic, model)
        sel = _sel(single)
        for launch in range(3):
            single.step(2)
            fusedsim.step(2, k=2, cfg=(0, chunk))
            _assert_equal(single, fusedsim, sel, f"{W}x{H} chunk {chunk} launch {launch}")


@pytest.mark.parametrize("storage,chunk", [("f32", 8), ("f16", 12), ("bf16", 24), ("f16", 16)])
def test_pair_kernel_interior_tiles(gpu, storage, chunk):
    """Tiles clear of the walls, the inlet and the outlet take the two-step kernel's interior form: obstacles
    inside such tiles, across their edges and touching their halo rows, bit for bit against single steps."""
    from lbm.sim import Simulation
    W, H = 700, 150
    mask = G.channel_mask(W, H, cylinder=(300, 70, 20))  # inside the x = 256..383 tiles
    mask[40:60, 380:395] = 1                               # across the x = 384 tile edge
    mask[chunk - 1:chunk + 1, 500:520] = 1                 # on the first interior tile row's halo
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.12, bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, 11),
              storage=storage)
    a, b = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    sel = _sel(a)
    for launch in range(4):
        a.step(2)
        b.step(2, k=2, cfg=(0, chunk))
        _assert_equal(a, b, sel, f"{storage} chunk {chunk} launch {launch}")


@pytest.mark.parametrize("seed", range(12))
def test_pair_kernel_random_configs(gpu, seed):
    """Seeded random lattices (130..900 x 20..260), random rectangular obstacles (some on tile edges and halo
    rows), either boundary mode, any storage and tile height: the two-step kernel's general, interior and
    all-fluid forms together equal two single steps bit for bit."""
    from lbm.sim import Simulation
    rng = np.random.default_rng(1000 + seed)
    W, H = int(rng.integers(130, 900)), int(rng.integers(20, 260))
    periodic = bool(rng.integers(0, 2))
    storage = ["f32", "f16", "bf16"][int(rng.integers(0, 3))]
    chunk = [8, 12, 16, 24][int(rng.integers(0, 4))]
    mask = G.periodic_mask(W, H) if periodic else G.channel_mask(W, H)
    for _ in range(int(rng.integers(0, 6))):
        h, w = int(rng.integers(1, 8)), int(rng.integers(1, 12))
        y, x = int(rng.integers(1, H - h)), int(rng.integers(2, W - w - 2))
        mask[y:y + h, x:x + w] = 1
    if periodic:
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=seed)
        kw = dict(bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy))
    else:
        prof = G.parabolic_profile(H, 0.05)
        kw = dict(bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed))
    kw.update(cs=0.12, storage=storage)
    a, b = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    sel = _sel(a)
    for launch in range(3):
        a.step(2)
        b.step(2, k=2, cfg=(0, chunk))
        _assert_equal(a, b, sel, f"{W}x{H} periodic={periodic} {storage} rows {chunk} launch {launch}")


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("clamp", [True, False])
def test_pair_kernel_clamp_forms(gpu, storage, clamp):
    """The two-step kernel takes the clamp setting as a template argument: with the clamps on, a state whose
    densities (0.2 .. 1.9) and speeds (up to 0.6) engage them; with them off, a milder state.  Both equal two
    single steps (run-time clamp flag) bit for bit, on a lattice with interior and edge tiles."""
    from lbm.sim import Simulation
    W, H = 300, 40
    rng = np.random.default_rng(21)
    lo, hi, um = (0.2, 1.9, 0.6) if clamp else (0.95, 1.05, 0.08)
    rho = rng.uniform(lo, hi, (H, W))
    ux, uy = rng.uniform(-um, um, (H, W)), rng.uniform(-um, um, (H, W))
    mask = G.periodic_mask(W, H)
    mask[15:20, 150:158] = 1
    kw = dict(cs=0.15, model="mrt", bc="periodic", mask=mask, f0=G.feq(rho, ux, uy), storage=storage,
              clamp=clamp)
    a, b = Simulation(W, H, 1.7, **kw), Simulation(W, H, 1.7, **kw)
    sel = _sel(a)
    for launch in range(3):
        a.step(2)
        b.step(2, k=2, cfg=(0, 8))
        _assert_equal(a, b, sel, f"clamp {clamp} launch {launch}")


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_pair_kernel_cylinder_large(gpu, storage):
    """4096 x 1024 cylinder (the benchmark lattice), the case's full 10000 steps with the 12-row two-step kernel
    (the benchmarked form): bitwise equal to single steps every 500 launches, the first 20 launches and the last."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    a = Simulation(case.W, case.H, case.tau_inv, storage=storage, **case.sim_kwargs())
    b = Simulation(case.W, case.H, case.tau_inv, storage=storage, **case.sim_kwargs())
    sel = _sel(a)
    n = case.steps // 2
    for launch in range(n):
        a.step(2)
        b.step(2, k=2, cfg=(0, 12))
        if launch < 20 or launch % 500 == 499 or launch == n - 1:
            _assert_equal(a, b, sel, f"launch {launch}")
    assert np.isfinite(b.diagnostics()).all()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_pair_kernel_16bit(gpu, storage):
    from lbm.sim import Simulation
    W, H = 301, 90
    mask = G.channel_mask(W, H, cylinder=(70, 45, 9))
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.12, bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, 5),
              storage=storage)
    a, b = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    sel = _sel(a)
    for launch in range(3):
        a.step(2)
        b.step(2, k=2, cfg=(0, 8))
        _assert_equal(a, b, sel, f"{storage} launch {launch}")


def test_pair_kernel_macros_and_limits(gpu):
    from lbm import _lib
    single, fusedsim = _pair(200, 64, False, "mrt")
    single.step(1)
    single.step(1, macros=True)
    fusedsim.step(2, k=2, cfg=(0, 16), macros=True)
    upd = _sel(single)[0][0]
    for a, b in ((single.rho, fusedsim.rho), (single.ux, fusedsim.ux), (single.uy, fusedsim.uy)):
        assert torch.equal(a[upd], b[upd])
    fusedsim.step(1)  # odd parity: the tile kernel does not apply
    with pytest.raises(_lib.LbmError):
        fusedsim.step(2, k=2, cfg=(0, 8))
    fusedsim.step(1)
    with pytest.raises(_lib.LbmError):
        fusedsim.step(6, k=6, cfg=(0, 8))  # tile kernels exist for k = 2 and k = 4 only
    with pytest.raises(_lib.LbmError):
        fusedsim.step(2, k=2, cfg=(0, 10))


@pytest.mark.parametrize("storage,cfg", [("f32", (0, 8)), ("f16", (0, 12)), ("bf16", (128, 32)), ("f32", None)])
def test_fused_n_equals_single_launches(gpu, storage, cfg):
    """lbm_step_fused_n / lbm_step_fused_h16_n (n launches from C, alternating buffers) equal n separate
    launches bit for bit; with an odd n the state ends in the second buffer."""
    import ctypes

    from lbm import _lib
    from lbm.sim import Simulation
    W, H = 301, 90
    mask = G.channel_mask(W, H, cylinder=(70, 45, 9))
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.12, bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, 9),
              storage=storage)
    a, b = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    k = 2 if cfg is None or cfg[0] == 0 else 4
    for _ in range(3):
        a._fused(k, False, cfg)  # one launch per call
    lib = _lib.load()
    fa, fb = b.f, torch.empty_like(b.f)
    c = None if cfg is None else ctypes.byref(_lib.LbmFusedCfg(*cfg))
    vp = ctypes.c_void_p
    st = vp(torch.cuda.current_stream().cuda_stream)
    if storage == "f32":
        rc = lib.lbm_step_fused_n(ctypes.byref(b.params), vp(b.mask.data_ptr()), vp(fa.data_ptr()), vp(fb.data_ptr()),
                                  vp(b.inlet_u.data_ptr()), 0, k, 3, c, st)
    else:
        rc = lib.lbm_step_fused_h16_n(ctypes.byref(b.params), vp(b.mask.data_ptr()), vp(fa.data_ptr()),
                                      vp(fb.data_ptr()), b.fmt, vp(b.inlet_u.data_ptr()), 0, k, 3, c, st)
    _lib.check(rc, "lbm_step_fused_n")
    torch.cuda.synchronize()
    sel = _sel(a)[(3 * k) & 1]
    assert int(((a.f != fb) & sel).sum()) == 0
    s2 = Simulation(W, H, 1.75, **kw)
    s2.step(3 * k, k=k, cfg=cfg)  # Simulation.step issues the run through the same entry point
    assert int(((a.f != s2.f) & sel).sum()) == 0 and s2.parity == a.parity


# ---------------------------------------------------------------------------------------------------------
# The four-step tile kernel (strip 0, k = 4, FP32 populations): bit for bit against four single steps
# ---------------------------------------------------------------------------------------------------------

QUAD_CHUNKS = [8, 12, 14, 16, 20, 24, 32]  # 128-wide tiles of 8 and 16 rows; 64-wide tiles of 12, 14, 20, 24, 32 rows


@pytest.mark.parametrize("chunk", QUAD_CHUNKS)
@pytest.mark.parametrize("periodic,model", [(False, "mrt"), (True, "mrt"), (False, "bgk"), (True, "bgk")])
def test_quad_kernel_equals_four_single_steps(gpu, chunk, periodic, model):
    """Ragged sizes, lattices narrower or lower than one 134 x (CH + 6) region, and periodic wraps."""
    sizes = [(333, 97), (134, 16), (70, 9)] if not periodic else [(300, 70), (100, 30), (40, 12), (66, 44)]
    for W, H in sizes:
        single, fusedsim = _pair(W, H, periodic, model)
        sel = _sel(single)
        for launch in range(3):
            single.step(4)
            fusedsim.step(4, k=4, cfg=(0, chunk))
            _assert_equal(single, fusedsim, sel, f"{W}x{H} chunk {chunk} launch {launch}")


@pytest.mark.parametrize("chunk", QUAD_CHUNKS)
def test_quad_kernel_interior_tiles(gpu, chunk):
    """Interior tiles (region clear of walls, inlet and outlet) with obstacles inside them, across tile edges
    and on the halo rows, and all-fluid interior tiles (the votes' fast path)."""
    from lbm.sim import Simulation
    W, H = 700, 150
    mask = G.channel_mask(W, H, cylinder=(300, 70, 20))
    mask[40:60, 380:395] = 1
    mask[chunk - 2:chunk + 1, 500:520] = 1
    prof = G.parabolic_profile(H, 0.05)
    kw = dict(cs=0.12, bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, 11))
    a, b = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    sel = _sel(a)
    for launch in range(4):
        a.step(4)
        b.step(4, k=4, cfg=(0, chunk))
        _assert_equal(a, b, sel, f"chunk {chunk} launch {launch}")


@pytest.mark.parametrize("seed", range(12))
def test_quad_kernel_random_configs(gpu, seed):
    """Seeded random lattices (130..900 x 20..260), random obstacles (some on tile edges and halo rows), either
    boundary mode, any tile height: the four-step kernel's forms together equal four single steps."""
    from lbm.sim import Simulation
    rng = np.random.default_rng(2000 + seed)
    W, H = int(rng.integers(130, 900)), int(rng.integers(20, 260))
    periodic = bool(rng.integers(0, 2))
    chunk = QUAD_CHUNKS[int(rng.integers(0, len(QUAD_CHUNKS)))]
    mask = G.periodic_mask(W, H) if periodic else G.channel_mask(W, H)
    for _ in range(int(rng.integers(0, 6))):
        h, w = int(rng.integers(1, 8)), int(rng.integers(1, 12))
        y, x = int(rng.integers(1, H - h)), int(rng.integers(2, W - w - 2))
        mask[y:y + h, x:x + w] = 1
    if periodic:
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=seed)
        kw = dict(bc="periodic", mask=mask, f0=G.feq(np.ones((H, W)), ux, uy))
    else:
        prof = G.parabolic_profile(H, 0.05)
        kw = dict(bc="channel", mask=mask, inlet_u=prof, f0=G.channel_initial_state(W, H, prof, 1e-3, seed))
    kw.update(cs=0.12)
    a, b = Simulation(W, H, 1.75, **kw), Simulation(W, H, 1.75, **kw)
    sel = _sel(a)
    for launch in range(3):
        a.step(4)
        b.step(4, k=4, cfg=(0, chunk))
        _assert_equal(a, b, sel, f"{W}x{H} periodic={periodic} rows {chunk} launch {launch}")


@pytest.mark.parametrize("clamp", [True, False])
def test_quad_kernel_clamp_forms(gpu, clamp):
    from lbm.sim import Simulation
    W, H = 300, 40
    rng = np.random.default_rng(22)
    lo, hi, um = (0.2, 1.9, 0.6) if clamp else (0.95, 1.05, 0.08)
    rho = rng.uniform(lo, hi, (H, W))
    ux, uy = rng.uniform(-um, um, (H, W)), rng.uniform(-um, um, (H, W))
    mask = G.periodic_mask(W, H)
    mask[15:20, 150:158] = 1
    kw = dict(cs=0.15, model="mrt", bc="periodic", mask=mask, f0=G.feq(rho, ux, uy), clamp=clamp)
    a, b = Simulation(W, H, 1.7, **kw), Simulation(W, H, 1.7, **kw)
    sel = _sel(a)
    for launch in range(3):
        a.step(4)
        b.step(4, k=4, cfg=(0, 8))
        _assert_equal(a, b, sel, f"clamp {clamp} launch {launch}")


def test_quad_kernel_macros_and_limits(gpu):
    from lbm import _lib
    single, fusedsim = _pair(200, 64, False, "mrt")
    single.step(3)
    single.step(1, macros=True)
    fusedsim.step(4, k=4, cfg=(0, 12), macros=True)
    upd = _sel(single)[0][0]
    for a, b in ((single.rho, fusedsim.rho), (single.ux, fusedsim.ux), (single.uy, fusedsim.uy)):
        assert torch.equal(a[upd], b[upd])
    with pytest.raises(_lib.LbmError):
        fusedsim.step(4, k=4, cfg=(0, 10))  # not in the four-step shape table
    fusedsim.step(1)
    with pytest.raises(_lib.LbmError):
        fusedsim.step(4, k=4, cfg=(0, 8))  # odd parity
    h16 = _pair(200, 64, False, "mrt")[0]
    from lbm.sim import Simulation
    s16 = Simulation(200, 64, 1.75, cs=0.12, mask=h16.mask_host, inlet_u=G.parabolic_profile(64, 0.05),
                     storage="f16")
    with pytest.raises(_lib.LbmError):
        s16.step(4, k=4, cfg=(0, 8))  # FP32 populations only


@pytest.mark.parametrize("chunk", [8, 24])
def test_quad_kernel_cylinder_large(gpu, chunk):
    """4096 x 1024 cylinder, the case's full 10000 steps with the four-step kernel (128 x 8 and 64 x 24 tiles):
    bitwise equal to single steps every 250 launches, the first 10 launches and the last."""
    from lbm.sim import Simulation
    case = configs.cylinder_large()
    a = Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs())
    b = Simulation(case.W, case.H, case.tau_inv, **case.sim_kwargs())
    sel = _sel(a)
    n = case.steps // 4
    for launch in range(n):
        a.step(4)
        b.step(4, k=4, cfg=(0, chunk))
        if launch < 10 or launch % 250 == 249 or launch == n - 1:
            _assert_equal(a, b, sel, f"launch {launch}")
    assert np.isfinite(b.diagnostics()).all()
