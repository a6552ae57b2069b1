"""Batched ensembles and captured graphs on the GPU: every member of an ensemble stepped by the persistent
kernel is bit-identical to the same lattice stepped on its own (lbm_step_aa / lbm_step_aa_h16), for
uneven and ragged sizes, both boundary modes, both models, all storage formats and any grid; every replay
of a captured graph equals the eager result."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm import geometry as G

pytestmark = pytest.mark.gpu

SIZES = [(64, 32), (70, 33), (129, 5), (256, 130), (384, 256), (128, 512), (3, 3), (200, 64), (512, 96)]


def _periodic_pair(sizes, storage="f32", model="mrt", cs=0.1, seed=11):
    from lbm.sim import Simulation
    a, b = [], []
    rng = np.random.default_rng(seed)
    for i, (W, H) in enumerate(sizes):
        ux, uy = G.random_phase_velocity(W, H, k0=min(8, max(1, min(W, H) // 4)), u_rms=0.04, seed=seed + i)
        f0 = G.feq(np.ones((H, W)), ux, uy)
        mask = G.periodic_mask(W, H)
        if W > 16 and H > 16:
            mask[H // 3:H // 3 + 5, W // 2:W // 2 + 4] = 1
        tau_inv = float(np.float32(G.nu_tau_inv(float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2)))))))
        kw = dict(cs=cs if model == "mrt" else 0.0, model=model, bc="periodic", mask=mask, f0=f0, storage=storage)
        a.append(Simulation(W, H, tau_inv, **kw))
        b.append(Simulation(W, H, tau_inv, **kw))
    return a, b


def _channel_pair(sizes, storage="f32", model="mrt", seed=21):
    from lbm.sim import Simulation
    a, b = [], []
    for i, (W, H) in enumerate(sizes):
        W, H = max(W, 8), max(H, 6)
        R = max(1, min(W, H) // 8)
        mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, R))
        prof = G.parabolic_profile(H, 0.05)
        f0 = G.channel_initial_state(W, H, prof, noise=1e-3, seed=seed + i)
        kw = dict(cs=0.12 if model == "mrt" else 0.0, model=model, bc="channel", mask=mask, inlet_u=prof, f0=f0,
                  storage=storage)
        a.append(Simulation(W, H, 1.7 - 0.01 * i, **kw))
        b.append(Simulation(W, H, 1.7 - 0.01 * i, **kw))
    return a, b


def _assert_same(ens_sims, solo_sims, what):
    torch.cuda.synchronize()
    for i, (s, r) in enumerate(zip(ens_sims, solo_sims)):
        assert s.parity == r.parity, f"{what}: member {i} parity"
        x, y = s.f.cpu().numpy(), r.f.cpu().numpy()
        bad = np.count_nonzero(x.view(np.uint8) != y.view(np.uint8))
        assert bad == 0, f"{what}: member {i} ({s.W}x{s.H}) differs in {bad} bytes"


def _solo_steps(sims, n):
    for s in sims:
        s.step(n)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_ensemble_periodic_bitwise(gpu, storage):
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES, storage=storage)
    ens = Ensemble(a)
    assert ens.cells == sum(W * H for W, H in SIZES)
    for n in (1, 2, 3):
        ens.step(n)
        _solo_steps(b, n)
        _assert_same(a, b, f"periodic {storage} after +{n}")
    ens.step_single(3)  # one lbm_ensemble_step launch per step
    _solo_steps(b, 3)
    _assert_same(a, b, f"periodic {storage} single launches")


@pytest.mark.parametrize("group_kb", [1, 300, 2000])
def test_ensemble_grouped_bitwise(gpu, group_kb):
    """Cache grouping (members advanced several steps per group) leaves every member bit-identical:
    1 KB puts each member in a group of its own, 300 KB packs a few, 2000 KB packs all."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES, seed=13)
    ens = Ensemble(a, group_bytes=group_kb << 10)
    for n in (5, 2, 1):
        ens.step(n)
        _solo_steps(b, n)
        _assert_same(a, b, f"group {group_kb} KB after +{n}")


@pytest.mark.parametrize("model", ["mrt", "bgk"])
def test_ensemble_channel_bitwise(gpu, model):
    from lbm.ensemble import Ensemble
    a, b = _channel_pair(SIZES, model=model)
    ens = Ensemble(a)
    ens.step(7)
    _solo_steps(b, 7)
    _assert_same(a, b, f"channel {model}")


@pytest.mark.parametrize("blocks", [8, 24, 4096])
def test_ensemble_any_grid(gpu, blocks):
    """Small grids (many tiles per workgroup) and a grid larger than the tile list give the same state."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES[:5], seed=31)
    ens = Ensemble(a, blocks=blocks)
    ens.step(4)
    _solo_steps(b, 4)
    _assert_same(a, b, f"blocks {blocks}")


def test_ensemble_single_member_and_solo_interleave(gpu):
    """One member; and single steps of the members between ensemble steps keep everything consistent."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair([(256, 128), (128, 64)], seed=41)
    ens = Ensemble(a)
    ens.step(1)
    _solo_steps(b, 1)
    ens.members[0].step(1)  # a member advanced alone changes its parity: the ensemble must refuse to mix
    with pytest.raises(ValueError):
        Ensemble(a)
    ens.members[1].step(1)
    ens.parity ^= 1
    _solo_steps(b, 1)
    ens.step(2)
    _solo_steps(b, 2)
    _assert_same(a, b, "interleaved")
    one, ref = _periodic_pair([(320, 200)], seed=43)
    e1 = Ensemble(one)
    e1.step(5)
    _solo_steps(ref, 5)
    _assert_same(one, ref, "single member")


@pytest.mark.parametrize("group_kb", [0, 300])
def test_ensemble_graph_replays(gpu, group_kb):
    """A captured 4-step graph, replayed 3 times with eager ensemble steps and synchronisations between
    the replays: every replay equals the eager result."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES, seed=51)
    ens = Ensemble(a, group_bytes=group_kb << 10)
    g = ens.capture(4)
    _assert_same(a, b, "capture records, does not run")
    for r in range(3):
        ens.replay(g)
        _solo_steps(b, 4)
        _assert_same(a, b, f"replay {r}")
        ens.step(2)  # eager work between replays
        _solo_steps(b, 2)
        torch.cuda.synchronize()
        _assert_same(a, b, f"eager after replay {r}")
    g.close()


def test_graph_of_single_lattice_steps(gpu):
    """The generic capture API records any liblbm launches: two single steps of one lattice + macros."""
    from lbm.ensemble import Graph
    a, b = _channel_pair([(300, 80)], seed=61)
    s, r = a[0], b[0]
    g = Graph()
    with g.capture():
        s.step(1, macros=True)
        s.step(1)
    s.parity, s.steps = r.parity, 0  # capture did not run anything
    _assert_same(a, b, "capture")
    for rep in range(3):
        g.replay()
        r.step(1, macros=True)
        r.step(1)
        _assert_same(a, b, f"replay {rep}")
        assert torch.equal(s.ux, r.ux) and torch.equal(s.rho, r.rho)
    g.close()


def test_ensemble_tuner(gpu, tmp_path, monkeypatch):
    from lbm.ensemble import Ensemble, tune_ensemble
    monkeypatch.setenv("LBM_TUNE_CACHE", str(tmp_path / "fused.json"))
    a, b = _periodic_pair(SIZES[:6], seed=71)
    ens = Ensemble(a)
    best = tune_ensemble(ens, steps=4, reps=1)
    assert best[0] % 8 == 0 and 8 <= best[0] <= ens.default_blocks and best[2] > 0
    assert (ens.blocks, ens.group_bytes) == (best[0], best[1])
    again = tune_ensemble(ens, steps=4, reps=1)  # from the cache
    assert tuple(again) == tuple(best)
    n = ens.members[0].steps
    _solo_steps(b, n)
    _assert_same(a, b, "after tuning")



# ------------------------------------------------------------------------ two-step (k = 2) ensemble path
def _observable(sim):
    p = sim.params
    op = O.params(p.W, p.H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc)
    return torch.from_numpy(O.observable(op, sim.mask_host, sim.parity, ghost=True)).to(sim.device)


def _assert_same_observable(ens_sims, solo_sims, what):
    """The two-step kernel leaves slots that no later step reads (the inlet's leaving slots) unwritten in
    its output buffer: compare every other population bit for bit."""
    torch.cuda.synchronize()
    for i, (s, r) in enumerate(zip(ens_sims, solo_sims)):
        assert s.parity == r.parity, f"{what}: member {i} parity"
        sel = _observable(r)
        bad = int(((s.f.view(torch.int16 if s.f.dtype == torch.int16 else torch.int32) !=
                    r.f.view(torch.int16 if r.f.dtype == torch.int16 else torch.int32)) & sel).sum())
        assert bad == 0, f"{what}: member {i} ({s.W}x{s.H}) differs in {bad} populations"


@pytest.mark.parametrize("storage,chunk", [("f32", 8), ("f32", 12), ("f32", 16), ("f32", 24), ("f16", 16),
                                           ("bf16", 8)])
def test_ensemble_two_step_periodic_bitwise(gpu, storage, chunk):
    """k = 2 launches over every member equal single steps of each member alone, on ragged sizes
    (including lattices narrower than one tile), mixed with k = 1 ensemble steps."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES, storage=storage, seed=81)
    ens = Ensemble(a, two_step=True, k=2, chunk=chunk)
    for n in (2, 4, 2):
        ens.step(n)
        _solo_steps(b, n)
        _assert_same_observable(a, b, f"{storage} rows {chunk} after +{n} (side {ens.side})")
    ens.step(1, k=1)  # odd parity: single steps only, in place in the current buffer
    _solo_steps(b, 1)
    with pytest.raises(ValueError):
        ens.step(2, k=2)
    ens.step(3, k=1)
    _solo_steps(b, 3)
    ens.step(2, k=2)
    _solo_steps(b, 2)
    _assert_same_observable(a, b, f"{storage} rows {chunk} mixed")


@pytest.mark.parametrize("model", ["mrt", "bgk"])
def test_ensemble_two_step_channel_bitwise(gpu, model):
    from lbm.ensemble import Ensemble
    a, b = _channel_pair(SIZES, model=model, seed=83)
    ens = Ensemble(a, two_step=True, k=2, chunk=12)
    ens.step(6)
    _solo_steps(b, 6)
    _assert_same_observable(a, b, f"channel {model}")
    for m in a:  # every member's own single steps continue from the ensemble's buffer
        m.step(2)
    _solo_steps(b, 2)
    _assert_same_observable(a, b, f"channel {model} then solo")


@pytest.mark.parametrize("blocks,group_kb", [(8, 0), (16, 300), (4096, 1), (0, 2000)])
def test_ensemble_two_step_grids_and_groups(gpu, blocks, group_kb):
    """Any grid (down to one workgroup per XCD) and any cache grouping give the same state."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES[:6], seed=85)
    ens = Ensemble(a, blocks=blocks, group_bytes=group_kb << 10, two_step=True, k=2, chunk=16)
    for n in (6, 2):
        ens.step(n)
        _solo_steps(b, n)
        _assert_same_observable(a, b, f"blocks {blocks} group {group_kb} KB after +{n}")


def test_ensemble_two_step_graph_replays(gpu):
    """A captured 8-step k = 2 graph returns to its side and parity; replayed 3 times with eager steps of
    both kinds between replays, every replay equals the eager single-step result."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES, seed=87)
    ens = Ensemble(a, two_step=True, k=2, chunk=16, group_bytes=300 << 10)
    with pytest.raises(ValueError):
        ens.capture(6)  # 3 two-step launches would end on the other buffer
    g = ens.capture(8)
    _assert_same_observable(a, b, "capture records, does not run")
    for r in range(3):
        ens.replay(g)
        _solo_steps(b, 8)
        _assert_same_observable(a, b, f"replay {r}")
        ens.step(2)  # moves the state to the other buffer: the graph must refuse to run
        _solo_steps(b, 2)
        with pytest.raises(RuntimeError):
            ens.replay(g)
        ens.step(2)
        ens.step(2, k=1)
        _solo_steps(b, 4)
        torch.cuda.synchronize()
        _assert_same_observable(a, b, f"eager after replay {r}")
    g.close()


def test_ensemble_two_step_limits(gpu):
    from lbm import _lib
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES[:3], seed=89)
    one = Ensemble(a)
    with pytest.raises(ValueError):
        one.step(2, k=2)  # built without second buffers
    one.close()
    ens = Ensemble(a, two_step=True)
    with pytest.raises(_lib.LbmError):
        ens.chunk = 10
        ens.step(2, k=2)
    ens.chunk = 16
    with pytest.raises(ValueError):
        ens.step(3, k=2)
    a[0].step(2, k=2, cfg=(0, 8))  # a member's own fused step swaps its buffers: the ensemble refuses
    with pytest.raises(RuntimeError):
        ens.step(2)


def test_ensemble_tuner_two_step(gpu, tmp_path, monkeypatch):
    from lbm.ensemble import Ensemble, tune_ensemble
    monkeypatch.setenv("LBM_TUNE_CACHE", str(tmp_path / "fused.json"))
    a, b = _periodic_pair(SIZES[:6], seed=91)
    ens = Ensemble(a, two_step=True)
    best = tune_ensemble(ens, steps=4, reps=1)
    assert best[3] in (1, 2, 4) and best[0] % 8 == 0 and best[2] > 0
    assert (ens.blocks, ens.group_bytes, ens.k, ens.chunk) == (best[0], best[1], best[3], best[4])
    assert tuple(tune_ensemble(ens, steps=4, reps=1)) == tuple(best)
    n = ens.members[0].steps
    _solo_steps(b, n)
    _assert_same_observable(a, b, "after tuning")


# ---------------------------------------------------------------------------------------------------------
# four-step launches, per-size-class plans and the 64-member workload
# ---------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("chunk", [8, 12, 20, 24])
@pytest.mark.parametrize("periodic", [True, False])
def test_ensemble_four_step_bitwise(gpu, chunk, periodic):
    """k = 4 launches (the four-step tile kernel over every member, FP32) equal single steps of each member
    alone, on ragged sizes including lattices narrower and lower than one tile region; mixed with k = 2 and
    k = 1 ensemble steps."""
    from lbm.ensemble import Ensemble
    a, b = _periodic_pair(SIZES, seed=91) if periodic else _channel_pair(SIZES, seed=92)
    ens = Ensemble(a, two_step=True, k=4, chunk=chunk)
    for n in (4, 8, 4):
        ens.step(n)
        _solo_steps(b, n)
        _assert_same_observable(a, b, f"k=4 rows {chunk} after +{n} (side {ens.side})")
    ens.chunk = 12
    ens.step(2, k=2)
    _solo_steps(b, 2)
    ens.chunk = chunk
    ens.step(1, k=1)
    _solo_steps(b, 1)
    with pytest.raises(ValueError):
        ens.step(4, k=4)  # odd parity
    ens.step(3, k=1)
    _solo_steps(b, 3)
    ens.step(4, k=4)
    _solo_steps(b, 4)
    _assert_same_observable(a, b, f"k=4 rows {chunk} mixed")


def test_ensemble_four_step_rejects_16bit(gpu):
    from lbm.ensemble import Ensemble
    a, _ = _periodic_pair(SIZES[:3], storage="f16", seed=93)
    ens = Ensemble(a, two_step=True)
    with pytest.raises(ValueError):
        ens.step(4, k=4)


def test_ensemble_class_plan_bitwise(gpu):
    """A plan giving each member range its own k and tile shape (k = 1 in place, k = 2, k = 4): members end on
    different buffer sides and every one equals its solo run; graphs of planned steps replay."""
    from lbm import _lib
    from lbm.ensemble import Ensemble, Segment
    sizes = SIZES + [(300, 200), (520, 300)]
    a, b = _periodic_pair(sizes, seed=94)
    ens = Ensemble(a, two_step=True)
    n = len(sizes)
    ens.plan = [Segment("x", 0, 3, 1, 0, 0, 0), Segment("y", 3, 6, 2, 16, 0, 1 << 20),
                Segment("z", 6, n, 4, 24, 64, 0)]
    ens.step(4)
    _solo_steps(b, 4)
    _assert_same_observable(a, b, "plan +4")
    assert ens.sides[:3] == [0] * 3 and ens.sides[3:6] == [0] * 3 and ens.sides[6:] == [1] * (n - 6)
    assert ens.side == -1
    with pytest.raises(_lib.LbmError):
        _lib.check(ens.lib.lbm_ensemble_side(ens.handle), "lbm_ensemble_side")  # the sides differ
    ens.step(8)
    _solo_steps(b, 8)
    _assert_same_observable(a, b, "plan +8")
    with pytest.raises(ValueError):
        ens.capture(4)  # k = 4 ranges need a multiple of 8 steps per replay
    g = ens.capture(8)
    for rep in range(3):
        ens.replay(g)
        _solo_steps(b, 8)
        _assert_same_observable(a, b, f"plan replay {rep}")
        ens.step(8)  # eager work between replays
        _solo_steps(b, 8)
    g.close()


def test_ensemble_class_tuner(gpu):
    """tune_classes gives every size class of a sorted member list its own shape, leaves the states where
    they were, and the planned steps stay bitwise equal to solo runs."""
    from lbm.ensemble import Ensemble, class_segments, sort_by_class, tune_classes
    sizes = sort_by_class([type("M", (), dict(W=w, H=h))() for w, h in
                           [(256, 256), (1024, 768), (384, 256), (512, 512), (1536, 1536), (768, 384)]])
    a, b = _periodic_pair([(m.W, m.H) for m in sizes], seed=95)
    ens = Ensemble(a, two_step=True)
    names = [sg.name for sg in class_segments(ens)]
    assert names == ["small", "medium", "large"]
    before = [m.f.clone() for m in a]
    plan = tune_classes(ens, steps=8, reps=1)
    assert [sg.name for sg in plan] == names and all(sg.k in (1, 2, 4) for sg in plan)
    for m, f in zip(a, before):
        assert torch.equal(m.f, f)
    ens.step(8)
    _solo_steps(b, 8)
    _assert_same_observable(a, b, "tuned plan")


def test_ensemble_workload_64_members(gpu):
    """The 64-member batched workload (configs.ensemble: sizes 256..2048 in both directions, nu log-uniform,
    random-phase fields): the persistent ensemble with k = 1, 2 and 4 launches equals every member's solo run
    bit for bit after 16 steps."""
    from lbm import configs
    from lbm.ensemble import Ensemble
    from lbm.sim import Simulation
    members = configs.ensemble()
    ens = Ensemble.from_members(members, two_step=True)
    solo = [Simulation(m.W, m.H, m.tau_inv, cs=m.cs, model="mrt", bc="periodic", f0=s.f.cpu().numpy())
            for m, s in zip(members, ens.members)]
    ens.chunk = 20
    ens.step(4, k=4)
    ens.chunk = 8
    ens.step(4, k=2)
    ens.step(4, k=1)
    ens.chunk = 24
    ens.step(4, k=4)
    for s in solo:
        s.step(16)
    _assert_same_observable(ens.members, solo, "64 members after 16 steps")
    ens.close()
