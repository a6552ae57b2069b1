"""Batched ensembles: many independent lattices advanced together by one persistent kernel per step.

The reference advances one lattice per process and replays a captured 2-launch CuPy graph
(lbm_smemaa_main.py:136-174).  Here ``Ensemble`` groups Simulation objects of any sizes (same model, bc,
storage format and parity), and ``lbm_ensemble_step`` advances all of them by one AA step in a single
launch whose workgroups walk a tile list spanning every member; ``lbm_ensemble_step_n`` advances groups of
members that fit in the Infinity Cache several steps at a time (members are independent, so the order in
which groups advance changes nothing).  With ``two_step=True`` every member also holds a second buffer and
``step(n, k=2)`` advances the ensemble two steps per launch with the two-step tile kernel, ``k=4`` four steps
with the four-step one (FP32 populations), moving each member's state between its two buffers.  A plan
(``tune_classes``) gives each size class -- a contiguous range of members, ``sort_by_class`` -- its own k, tile
rows, grid and cache grouping (``lbm_ensemble_run``).  ``Graph`` records any sequence of liblbm launches on a
side stream (``capture``) and replays it; ``Ensemble.capture(n, k)`` records n ensemble steps.  Every member's
state is bit-identical to what its own ``Simulation.step`` would produce.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass

import torch

from . import _lib
from .sim import Simulation, _stream


class LbmInstance(ctypes.Structure):
    """Mirror of ``lbm_instance`` (include/lbm.h)."""

    _fields_ = [("p", _lib.LbmParams), ("mask", ctypes.c_void_p), ("f", ctypes.c_void_p),
                ("inlet_u", ctypes.c_void_p), ("f_alt", ctypes.c_void_p)]


# Size classes of ensemble members, by cell count: the tuner gives each its own launch shape.  Small members
# (fewer than 2^19 cells) have few interior tiles, large ones (2^21 and more) fill the GPU on their own.
CLASS_BOUNDS = (1 << 19, 1 << 21)
CLASS_NAMES = ("small", "medium", "large")


def size_class(W, H):
    """Size class name of a W x H member (CLASS_BOUNDS on its cell count)."""
    cells = int(W) * int(H)
    return CLASS_NAMES[sum(cells >= b for b in CLASS_BOUNDS)]


def sort_by_class(members):
    """Members (anything with W and H) in size-class order, stable within a class, so that each class is a
    contiguous range of the ensemble."""
    return sorted(members, key=lambda m: CLASS_NAMES.index(size_class(m.W, m.H)))


@dataclass
class Segment:
    """One planned range of an ensemble: members [m0, m1) advanced with k steps per launch, `chunk` tile rows,
    `blocks` persistent workgroups (0 = the kernel's resident grid) and cache groups of `group_bytes`."""
    name: str
    m0: int
    m1: int
    k: int = 1
    chunk: int = 0
    blocks: int = 0
    group_bytes: int = 0
    us_per_step: float = 0.0


class Graph:
    """A recorded sequence of liblbm launches (lbm_graph_*).  Use as a context manager to capture:

        g = Graph()
        with g.capture():
            sim.step(2)
        g.replay()      # runs both steps again (the state advances; parity bookkeeping is the caller's)
    """

    def __init__(self):
        self.lib = _lib.load()
        self.handle = ctypes.c_void_p()
        self.stream = torch.cuda.Stream()  # default priority; capture needs a non-default stream

    def capture(self):
        g = self

        class _Cap:
            def __enter__(self_):
                if g.handle:
                    raise RuntimeError("graph already captured")
                torch.cuda.current_stream().synchronize()
                self_.ctx = torch.cuda.stream(g.stream)
                self_.ctx.__enter__()
                _lib.check(g.lib.lbm_graph_begin(ctypes.c_void_p(g.stream.cuda_stream)), "lbm_graph_begin")
                return g

            def __exit__(self_, et, ev, tb):
                rc = g.lib.lbm_graph_end(ctypes.c_void_p(g.stream.cuda_stream), ctypes.byref(g.handle))
                self_.ctx.__exit__(et, ev, tb)
                if et is None:
                    _lib.check(rc, "lbm_graph_end")
                return False

        return _Cap()

    def replay(self, stream=None):
        """Launch the recorded sequence on `stream` (default: the current stream)."""
        if not self.handle:
            raise RuntimeError("graph not captured")
        st = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
        _lib.check(self.lib.lbm_graph_launch(self.handle, st), "lbm_graph_launch")

    def close(self):
        if self.handle:
            self.lib.lbm_graph_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Independent lattices stepped together.  ``members`` are Simulation objects sharing model, bc,
    storage format and parity; their buffers must not be replaced while the ensemble exists (the
    single-step path updates them in place, so ``Simulation.step(k=1)`` keeps them valid too).

    ``two_step=True`` gives every member a second buffer (``Simulation.f_alt``, allocated if missing) for
    ``step(n, k=2)`` and ``step(n, k=4)``; the ensemble then swaps each member's ``f`` and ``f_alt`` so that
    ``f`` always holds the state.  ``k`` and ``chunk`` (tile rows: 8/12/16/24 for k = 2; 8/12/20/24 for k = 4, FP32
    populations) are the defaults of ``step``.

    ``plan`` (set by ``tune_classes``) gives contiguous member ranges -- the size classes, see
    ``sort_by_class`` -- their own launch shape: a list of ``Segment``; ``step(n)`` without an explicit k then
    advances each range with its own k, tile rows, grid and cache grouping (lbm_ensemble_run)."""

    def __init__(self, members, blocks=0, group_bytes=0, two_step=False, k=1, chunk=16):
        self.lib = _lib.load()
        self.members = list(members)
        if not self.members:
            raise ValueError("an ensemble needs at least one member")
        m0 = self.members[0]
        for m in self.members:
            if (m.model, m.bc, m.fmt, m.parity) != (m0.model, m0.bc, m0.fmt, m0.parity):
                raise ValueError("ensemble members must share model, bc, storage format and parity")
        self.fmt = m0.fmt
        self.parity = m0.parity
        self.blocks = int(blocks)
        self.group_bytes = int(group_bytes)
        self.two_step = bool(two_step)
        self.k, self.chunk = int(k), int(chunk)
        self.plan = None
        if self.two_step:
            for m in self.members:
                if m.f_alt is None:
                    m.f_alt = torch.empty_like(m.f)
        self._table = (LbmInstance * len(self.members))()
        for i, m in enumerate(self.members):
            self._table[i] = LbmInstance(m.params, m.mask.data_ptr(), m.f.data_ptr(), m.inlet_u.data_ptr(),
                                         m.f_alt.data_ptr() if self.two_step else None)
        # the buffers as registered: buffer 0 holds the state at side 0, buffer 1 at side 1
        self._bufs = [(m.f.data_ptr(), m.f_alt.data_ptr() if self.two_step else None) for m in self.members]
        self.sides = [0] * len(self.members)
        self.handle = ctypes.c_void_p()
        _lib.check(self.lib.lbm_ensemble_create(self._table, len(self.members), self.fmt, ctypes.byref(self.handle)),
                   "lbm_ensemble_create")
        tiles, cells, blk = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
        _lib.check(self.lib.lbm_ensemble_info(self.handle, ctypes.byref(tiles), ctypes.byref(cells),
                                              ctypes.byref(blk)), "lbm_ensemble_info")
        self.total_tiles, self.cells, self.default_blocks = tiles.value, cells.value, blk.value

    @property
    def side(self):
        """The buffer holding every member's state (0: f, 1: f_alt), or -1 when members differ."""
        return self.sides[0] if all(v == self.sides[0] for v in self.sides) else -1

    @classmethod
    def from_members(cls, members, storage="f32", device="cuda", blocks=0, workers=8, two_step=False,
                     by_class=False, arith="exact"):
        """Build the Simulations of ``configs.ensemble`` members (periodic, MRT + LES) and group them (in size-class
        order with ``by_class``, see ``sort_by_class``).  The initial states are computed on `workers` host
        threads, at most `workers` of them held at once.  ``arith="fast"``: the hardware reciprocal and square
        root in every member's collision (lbm_params.fast)."""
        import concurrent.futures as cf
        import numpy as np
        members = list(members)
        if by_class:
            members = sort_by_class(members)
        sims = []
        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            for i in range(0, len(members), workers):
                batch = members[i:i + workers]
                for m, f0 in zip(batch, ex.map(lambda mm: mm.f0(np.float32), batch)):
                    sims.append(Simulation(m.W, m.H, m.tau_inv, cs=m.cs, model="mrt", bc="periodic", f0=f0,
                                           device=device, storage=storage, arith=arith))
        return cls(sims, blocks=blocks, two_step=two_step)

    def pair_blocks(self, chunk, k=2):
        """Workgroups of one k-step launch (k = 2 or 4) with `chunk` tile rows over all members: one per tile."""
        tiles, blk = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.lbm_ensemble_run_info(self.handle, int(k), int(chunk), ctypes.byref(tiles),
                                                  ctypes.byref(blk)), "lbm_ensemble_run_info")
        return blk.value

    def _check_buffers(self):
        for m, b, sd in zip(self.members, self._bufs, self.sides):
            if m.f.data_ptr() != b[sd] or (self.two_step and m.f_alt.data_ptr() != b[sd ^ 1]):
                raise RuntimeError("a member's state buffer moved (its own fused steps swap buffers): "
                                   "rebuild the ensemble")

    def _segments(self, k):
        """The (m0, m1, k, chunk, blocks, group_bytes) launches of one step call: the plan when no k is forced,
        else every member with k and the ensemble's defaults."""
        if k is None and self.plan is not None:
            return [(sg.m0, sg.m1, sg.k, sg.chunk, sg.blocks, sg.group_bytes) for sg in self.plan]
        k = self.k if k is None else int(k)
        return [(0, len(self.members), k, self.chunk, self.blocks, self.group_bytes)]

    def _check_k(self, n, k):
        if k not in (1, 2, 4):
            raise ValueError("ensemble steps take k = 1, 2 or 4")
        if k > 1:
            if not self.two_step:
                raise ValueError(f"k = {k} needs an ensemble built with two_step=True")
            if self.parity != 0 or n % k:
                raise ValueError(f"k = {k} advances a multiple of {k} steps from an even state")
            if k == 4 and self.fmt != _lib.FMT_F32:
                raise ValueError("k = 4 takes FP32 populations only")

    def _run(self, m0, m1, k, chunk, n, blocks, group_bytes, stream):
        _lib.check(self.lib.lbm_ensemble_run(self.handle, m0, m1, k, chunk if k > 1 else 0, n, self.parity,
                                             group_bytes, blocks, stream), "lbm_ensemble_run")
        if k > 1 and (n // k) & 1:
            for i in range(m0, m1):
                m = self.members[i]
                m.f, m.f_alt = m.f_alt, m.f
                self.sides[i] ^= 1

    def _launch(self, parity, n, stream, k=None):
        assert parity == self.parity
        for m0, m1, kk, ch, blocks, gb in self._segments(k):
            self._run(m0, m1, kk, ch, n, blocks, gb, stream)

    def step(self, n=1, k=None):
        """n AA steps of every member: members grouped by ``group_bytes`` (0 = all in one group), each
        group advanced n steps (one persistent launch per step, or per k steps with k = 2 or 4) before the
        next; with a plan and no k, each planned range with its own shape."""
        self._check_buffers()
        n = int(n)
        if n > 0:
            for _, _, kk, _, _, _ in self._segments(k):
                self._check_k(n, kk)
            self._launch(self.parity, n, _stream(), k)
            self.parity = (self.parity + n) & 1
            self._sync_members(n)
        return self

    def step_single(self, n=1):
        """n steps as n calls of lbm_ensemble_step (one launch over all members per step)."""
        self._check_buffers()
        st = _stream()
        for _ in range(int(n)):
            _lib.check(self.lib.lbm_ensemble_step(self.handle, self.parity, self.blocks, st), "lbm_ensemble_step")
            self.parity ^= 1
        self._sync_members(int(n))
        return self

    def step_separate(self, n=1, lattice_major=False):
        """The same n steps as one launch per member per step (for comparison): step-major (every member
        one step, n times) or lattice-major (each member its n steps in turn)."""
        if lattice_major:
            for m in self.members:
                m.step(int(n))
        else:
            for _ in range(int(n)):
                for m in self.members:
                    m.step(1)
        self.parity = (self.parity + int(n)) & 1
        return self

    def _sync_members(self, n):
        for m in self.members:
            m.parity = self.parity
            m.steps += n

    def capture(self, n=2, k=None):
        """A Graph that performs n ensemble steps from the current parity and sides.  Each replay must return
        to both, so n is a multiple of 2k for every k in use; replay it with ``replay(g)``."""
        n = int(n)
        segs = self._segments(k)
        for _, _, kk, _, _, _ in segs:
            if n < 2 or n % (2 * kk):
                raise ValueError(f"capture a multiple of {2 * kk} steps with k = {kk}")
            self._check_k(n, kk)
        self._check_buffers()
        g = Graph()
        sides = list(self.sides)
        with g.capture():
            self._launch(self.parity, n, ctypes.c_void_p(g.stream.cuda_stream), k)
        assert self.sides == sides  # an even number of launches per member
        g.steps, g.at = n, (self.parity, tuple(self.sides))
        return g

    def replay(self, g, times=1):
        """Replay a graph from ``capture``; the ensemble must be at the parity and side it was captured at
        (the launches' buffers and parities are fixed in the graph)."""
        if (self.parity, tuple(self.sides)) != g.at:
            raise RuntimeError(f"graph captured at (parity, sides) {g.at[0], g.at[1][:4]}..., ensemble at "
                               f"{(self.parity, tuple(self.sides)[:4])}...")
        self._check_buffers()
        for _ in range(int(times)):
            g.replay()
            self._sync_members(g.steps)

    def close(self):
        if self.handle:
            self.lib.lbm_ensemble_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------------- tuning
_CACHE_ENV = "LBM_TUNE_CACHE"


def _cache_path():
    base = os.environ.get(_CACHE_ENV, os.path.join(os.path.expanduser("~"), ".cache", "lbm-mi355x", "fused.json"))
    return os.path.join(os.path.dirname(base), "ensemble.json")


def size_classes(ens):
    """Histogram of member sizes: the tuning key."""
    hist = {}
    for m in ens.members:
        hist[f"{m.W}x{m.H}"] = hist.get(f"{m.W}x{m.H}", 0) + 1
    return ",".join(f"{k}:{v}" for k, v in sorted(hist.items()))


def time_ensemble(ens, blocks, group_bytes, steps=20, reps=2, k=1, chunk=16):
    """Mean microseconds per ensemble step with grid `blocks`, grouping `group_bytes` and `k` steps per
    launch (tile rows `chunk` for k = 2), in calls of `steps` steps (the state advances by
    steps * (reps + 1))."""
    old = ens.blocks, ens.group_bytes, ens.chunk
    ens.blocks, ens.group_bytes, ens.chunk = int(blocks), int(group_bytes), int(chunk)
    try:
        ens.step(steps, k=k)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ens.step(steps, k=k)
        e1.record()
        torch.cuda.synchronize()
    finally:
        ens.blocks, ens.group_bytes, ens.chunk = old
    return e0.elapsed_time(e1) * 1e3 / (steps * reps)


GROUPS_MB = (0, 32, 64, 128, 192)
PAIR_CHUNKS = (8, 12, 16, 24)


QUAD_CHUNKS = (8, 12, 20, 24)


def tune_candidates(ens, steps):
    """(k, chunk, blocks) launch shapes to time: the single-step kernel at 2, 4, 6 or 8 workgroups per CU
    (never more than are resident at once), and, for a two-step ensemble at an even state, the two-step kernel
    (even `steps`) and the four-step kernel (FP32, `steps` a multiple of 4) at each tile shape (one workgroup per
    tile: blocks 0)."""
    full = ens.default_blocks
    cands = [(1, 0, b) for b in sorted({max(8, full * w // 8 // 8 * 8) for w in (2, 4, 6, 8)})]
    if ens.two_step and ens.parity == 0 and steps % 2 == 0:
        cands += [(2, ch, 0) for ch in PAIR_CHUNKS]
    if ens.two_step and ens.parity == 0 and steps % 4 == 0 and ens.fmt == _lib.FMT_F32:
        cands += [(4, ch, 0) for ch in QUAD_CHUNKS]
    return cands


def tune_ensemble(ens, use_cache=True, steps=20, reps=2, verbose=False):
    """Pick the launch shape (k, tile rows, persistent grid) and the cache grouping (GROUPS_MB, 0 = none)
    for this mix of size classes and `steps` per call; sets ens.k, ens.chunk, ens.blocks and
    ens.group_bytes and returns (blocks, group_bytes, us_per_step, k, chunk)."""
    dev = torch.cuda.get_device_name(ens.members[0].device)
    m0 = ens.members[0]
    two = ens.two_step and ens.parity == 0 and steps % 2 == 0
    key = f"{dev}|fmt{ens.fmt}|model{m0.model}|bc{m0.bc}|steps{steps}|k{2 if two else 1}|v3|{size_classes(ens)}"
    cache = {}
    if use_cache:
        try:
            with open(_cache_path()) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        best = tuple(cache[key])
    else:
        best = None
        ens.tune_table = []
        for gmb in GROUPS_MB:
            for k, ch, b in tune_candidates(ens, steps):
                us = time_ensemble(ens, b, gmb << 20, steps=steps, reps=reps, k=k, chunk=ch or ens.chunk)
                ens.tune_table.append((k, ch, b, gmb, us))
                if verbose:
                    print(f"  k={k} rows={ch} blocks={b} group={gmb} MB: {us:.1f} us/step", flush=True)
                if best is None or us < best[2]:
                    best = (b, gmb << 20, us, k, ch or ens.chunk)
        if use_cache:
            cache[key] = list(best)
            try:
                os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
                with open(_cache_path() + ".tmp", "w") as fh:
                    json.dump(cache, fh, indent=1, sort_keys=True)
                os.replace(_cache_path() + ".tmp", _cache_path())
            except OSError:
                pass
    ens.blocks, ens.group_bytes, ens.k, ens.chunk = int(best[0]), int(best[1]), int(best[3]), int(best[4])
    return best


def class_segments(ens):
    """The ensemble's contiguous runs of members of one size class, as Segments (an ensemble built from
    ``sort_by_class`` members has one per class present)."""
    segs, m0 = [], 0
    names = [size_class(m.W, m.H) for m in ens.members]
    for i in range(1, len(names) + 1):
        if i == len(names) or names[i] != names[m0]:
            segs.append(Segment(names[m0], m0, i))
            m0 = i
    return segs


def _time_segment(ens, sg, k, chunk, blocks, group_bytes, steps, reps):
    """Mean microseconds per step of the members of `sg` alone with the given shape.  `steps` is a multiple of
    8, so every run makes an even number of launches for any k and leaves each member's state in the buffer it
    started in, at the same parity."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    st = _stream()
    ens._run(sg.m0, sg.m1, k, chunk, steps, blocks, group_bytes, st)
    e0.record()
    for _ in range(reps):
        ens._run(sg.m0, sg.m1, k, chunk, steps, blocks, group_bytes, st)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / (steps * reps)


def tune_classes(ens, steps=24, reps=2, verbose=False):
    """A per-size-class plan: for each class range, the fastest (k, tile rows, grid, cache grouping) over
    ``tune_candidates`` x GROUPS_MB, timed on that range alone.  Sets ``ens.plan`` and returns it.  Timing
    advances the classes by different numbers of steps, so the states are copied first and put back afterwards:
    the ensemble ends where it started."""
    if ens.parity != 0 or steps % 8:
        raise ValueError("tune classes from an even state with steps a multiple of 8")
    ens._check_buffers()
    saved = [m.f.clone() for m in ens.members]
    sides = list(ens.sides)
    plan = []
    try:
        for sg in class_segments(ens):
            best = None
            for gmb in GROUPS_MB:
                for k, ch, b in tune_candidates(ens, steps):
                    us = _time_segment(ens, sg, k, ch, b, gmb << 20, steps, reps)
                    if verbose:
                        print(f"  {sg.name}: k={k} rows={ch} blocks={b} group={gmb} MB: {us:.1f} us/step", flush=True)
                    if best is None or us < best.us_per_step:
                        best = Segment(sg.name, sg.m0, sg.m1, k, ch, b, gmb << 20, us)
            plan.append(best)
    finally:
        assert ens.sides == sides  # even launch counts: every state is back in its own buffer
        for m, f in zip(ens.members, saved):
            m.f.copy_(f)
        torch.cuda.synchronize()
    ens.plan = plan
    return plan
