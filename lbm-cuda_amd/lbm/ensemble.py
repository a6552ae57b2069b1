"""Batched ensembles: many independent lattices advanced together by one persistent kernel per step.

The reference advances one lattice per process and replays a captured 2-launch CuPy graph
(lbm_smemaa_main.py:136-174).  Here ``Ensemble`` groups Simulation objects of any sizes (same model, bc,
storage format and parity), and ``lbm_ensemble_step`` advances all of them by one AA step in a single
launch whose workgroups walk a tile list spanning every member; ``lbm_ensemble_step_n`` advances groups of
members that fit in the Infinity Cache several steps at a time (members are independent, so the order in
which groups advance changes nothing).  ``Graph`` records any sequence of liblbm
launches on a side stream (``capture``) and replays it; ``Ensemble.capture(n)`` records n ensemble steps.
Every member's state is bit-identical to what its own ``Simulation.step`` would produce.
"""
from __future__ import annotations

import ctypes
import json
import os

import torch

from . import _lib
from .sim import Simulation, _stream


class LbmInstance(ctypes.Structure):
    """Mirror of ``lbm_instance`` (include/lbm.h)."""

    _fields_ = [("p", _lib.LbmParams), ("mask", ctypes.c_void_p), ("f", ctypes.c_void_p),
                ("inlet_u", ctypes.c_void_p)]


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
    single-step path updates them in place, so ``Simulation.step(k=1)`` keeps them valid too)."""

    def __init__(self, members, blocks=0, group_bytes=0):
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
        self._table = (LbmInstance * len(self.members))()
        for i, m in enumerate(self.members):
            self._table[i] = LbmInstance(m.params, m.mask.data_ptr(), m.f.data_ptr(), m.inlet_u.data_ptr())
        self._ptrs = [m.f.data_ptr() for m in self.members]
        self.handle = ctypes.c_void_p()
        _lib.check(self.lib.lbm_ensemble_create(self._table, len(self.members), self.fmt, ctypes.byref(self.handle)),
                   "lbm_ensemble_create")
        tiles, cells, blk = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
        _lib.check(self.lib.lbm_ensemble_info(self.handle, ctypes.byref(tiles), ctypes.byref(cells),
                                              ctypes.byref(blk)), "lbm_ensemble_info")
        self.total_tiles, self.cells, self.default_blocks = tiles.value, cells.value, blk.value

    @classmethod
    def from_members(cls, members, storage="f32", device="cuda", blocks=0, workers=8):
        """Build the Simulations of ``configs.ensemble`` members (periodic, MRT + LES) and group them.  The
        initial states are computed on `workers` host threads, at most `workers` of them held at once."""
        import concurrent.futures as cf
        import numpy as np
        members = list(members)
        sims = []
        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            for i in range(0, len(members), workers):
                batch = members[i:i + workers]
                for m, f0 in zip(batch, ex.map(lambda mm: mm.f0(np.float32), batch)):
                    sims.append(Simulation(m.W, m.H, m.tau_inv, cs=m.cs, model="mrt", bc="periodic", f0=f0,
                                           device=device, storage=storage))
        return cls(sims, blocks=blocks)

    def _check_buffers(self):
        if [m.f.data_ptr() for m in self.members] != self._ptrs:
            raise RuntimeError("a member's state buffer moved (fused steps swap buffers): rebuild the ensemble")

    def _launch(self, parity, n, stream):
        _lib.check(self.lib.lbm_ensemble_step_n(self.handle, parity, n, self.group_bytes, self.blocks, stream),
                   "lbm_ensemble_step_n")

    def step(self, n=1):
        """n AA steps of every member: members grouped by ``group_bytes`` (0 = all in one group), each
        group advanced n steps (one persistent launch per step) before the next."""
        self._check_buffers()
        n = int(n)
        if n > 0:
            self._launch(self.parity, n, _stream())
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

    def capture(self, n=2):
        """A Graph that performs n ensemble steps from the current parity.  n must be even so that each
        replay returns to the same parity; call ``replayed(g)`` after each replay for the bookkeeping."""
        n = int(n)
        if n < 2 or n % 2:
            raise ValueError("capture an even number of steps")
        self._check_buffers()
        g = Graph()
        with g.capture():
            self._launch(self.parity, n, ctypes.c_void_p(g.stream.cuda_stream))
        g.steps = n
        return g

    def replay(self, g, times=1):
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


def time_ensemble(ens, blocks, group_bytes, steps=20, reps=2):
    """Mean microseconds per ensemble step with grid `blocks` and grouping `group_bytes`, in calls of
    `steps` steps (the state advances by steps * (reps + 1))."""
    old = ens.blocks, ens.group_bytes
    ens.blocks, ens.group_bytes = int(blocks), int(group_bytes)
    ens.step(steps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ens.step(steps)
    e1.record()
    torch.cuda.synchronize()
    ens.blocks, ens.group_bytes = old
    return e0.elapsed_time(e1) * 1e3 / (steps * reps)


GROUPS_MB = (0, 32, 64, 128, 192)


def tune_ensemble(ens, use_cache=True, steps=20, reps=2, verbose=False):
    """Pick the persistent grid (2, 4, 6 or 8 workgroups per CU, never more than are resident at once) and
    the cache grouping (GROUPS_MB, 0 = none) for this mix of size classes and `steps` per call; sets
    ens.blocks and ens.group_bytes and returns (blocks, group_bytes, us_per_step)."""
    dev = torch.cuda.get_device_name(ens.members[0].device)
    m0 = ens.members[0]
    key = f"{dev}|fmt{ens.fmt}|model{m0.model}|bc{m0.bc}|steps{steps}|{size_classes(ens)}"
    cache = {}
    if use_cache:
        try:
            with open(_cache_path()) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        ens.blocks, ens.group_bytes = int(cache[key][0]), int(cache[key][1])
        return tuple(cache[key])
    full = ens.default_blocks
    cands = sorted({max(8, full * w // 8 // 8 * 8) for w in (2, 4, 6, 8)})
    best = None
    for gmb in GROUPS_MB:
        for b in cands:
            us = time_ensemble(ens, b, gmb << 20, steps=steps, reps=reps)
            if verbose:
                print(f"  blocks={b} group={gmb} MB: {us:.1f} us/step", flush=True)
            if best is None or us < best[2]:
                best = (b, gmb << 20, us)
    ens.blocks, ens.group_bytes = best[0], best[1]
    if use_cache:
        cache[key] = list(best)
        try:
            os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
            with open(_cache_path() + ".tmp", "w") as fh:
                json.dump(cache, fh, indent=1, sort_keys=True)
            os.replace(_cache_path() + ".tmp", _cache_path())
        except OSError:
            pass
    return best
