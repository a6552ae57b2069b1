"""Autotuner for the temporally blocked kernel: picks the strip width, the chunk height and k for a lattice.

The reference fixes its tile (SQUARE_A = 24) and T at compile time and mirrors them by hand in the driver
(lbm_smemaa_main.py:93-95) -- its compile-time traversal-order table (lbm_core_smemaa_policy.cu:12-24) is
the only tuning knob.  Here every candidate is timed on the GPU with HIP events on scratch copies of the
state (the simulation itself is not advanced) and the winner is cached per lattice size class.  From an even
state the candidates are the tile kernels (strip 0): the two-step kernel for k = 2 and, for FP32 populations, the
four-step kernel for k = 4; the strip sweep is offered from odd states only.
"""
from __future__ import annotations

import ctypes
import json
import os

import torch

from . import _lib

STRIPS = (64, 128)
CHUNKS = (16, 32, 64, 128, 256)
KS = (2, 4, 6, 8)

_CACHE_ENV = "LBM_TUNE_CACHE"


def _cache_path():
    return os.environ.get(_CACHE_ENV, os.path.join(os.path.expanduser("~"), ".cache", "lbm-mi355x", "fused.json"))


def _key(sim, k):
    p = sim.params
    dev = torch.cuda.get_device_name(sim.device)
    rb = int(bool(p.row_base)) and sim.fmt != _lib.FMT_F32
    win = "" if not rb or not p.window else "|win{0}x{1}+{2}+{3}".format(*(getattr(p.window.contents, a)
                                                                          for a in ("w", "h", "x0", "y0")))
    return f"{dev}|{p.W}x{p.H}|bc{p.bc}|model{p.model}|les{int(p.cs != 0)}|fmt{sim.fmt}|rb{rb}{win}|fast{p.fast}|k{k}"


def load_cache():
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(cache):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
        os.replace(path + ".tmp", path)
    except OSError:
        pass  # a read-only home is not an error: tuning still returns its result


def time_config(sim, k, cfg, reps=10, bufs=None):
    """Microseconds per collide-stream step of lbm_step_fused_cfg with `cfg` (strip, chunk): `reps` untimed
    launches (clocks settle after the previous candidate), then the faster of two timed batches of `reps`."""
    L = sim.lib
    if bufs is None:
        bufs = (sim.f.clone(), torch.empty_like(sim.f))
    a, b = bufs
    c = None if cfg is None or cfg[0] is None else _lib.LbmFusedCfg(int(cfg[0]), int(cfg[1]))
    cp = None if c is None else ctypes.byref(c)  # (None, None): no configuration, the library's choice
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    par = sim.parity

    def launch(src, dst, parity):
        ptrs = (ctypes.c_void_p(sim.mask.data_ptr()), ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()))
        if sim.fmt == _lib.FMT_F32:
            rc = L.lbm_step_fused_cfg(ctypes.byref(sim.params), *ptrs, ctypes.c_void_p(sim.inlet_u.data_ptr()), parity,
                                      k, cp, None, None, None, st)
        else:
            rc = L.lbm_step_fused_h16(ctypes.byref(sim.params), *ptrs, sim.fmt, ctypes.c_void_p(sim.inlet_u.data_ptr()),
                                      parity, k, cp, None, None, None, st)
        _lib.check(rc, "lbm_step_fused")

    reps += reps & 1  # an even number of launches returns the state to buffer a and to parity par

    def batch():
        for r in range(reps):
            src, dst = (a, b) if r % 2 == 0 else (b, a)
            launch(src, dst, (par + r * k) & 1)

    batch()  # warm-up
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = None
    for _ in range(2):
        e0.record()
        batch()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / (reps * k)
        best = us if best is None else min(best, us)
    return best


def p_fast(sim):
    return bool(sim.params.fast)


def candidates(sim, k):
    """(strip, chunk) pairs to time.  From an even state: the tile kernels only (strip 0, chunk = tile rows) --
    the two-step kernel for k = 2, the four-step kernel for k = 4 (FP32 populations) -- and nothing for other k:
    on 4096 x 1024 the strip sweep ran 61-85 us per step against 20-36 us for the tile kernels and 34-45 us for
    single steps, so it cannot win there.  From an odd state (no tile kernel applies): (None, None) -- no
    configuration, where the library takes one in-place single step or a copy and then a tile launch and single
    steps (64 us per step at k = 4 on 4096 x 1024 through lbm_step_fused_cfg, 42 us through lbm_step_fused_n) --
    and the strip-sweep shapes (80 us per step at k = 4 there), which the fast arithmetic and row-based 16-bit
    storage do not have."""
    H = sim.params.H
    if sim.fmt != _lib.FMT_F32 and sim.params.row_base:  # row-based 16-bit storage: the two-step kernel only
        if sim.parity == 1:
            return [(None, None)]
        return [(0, ch) for ch in (8, 12, 16, 24)] if k == 2 else []
    if p_fast(sim) and sim.parity == 1:  # fast arithmetic: the tile kernels only
        return [(None, None)]
    if sim.parity == 0:
        if k == 2:
            return [(0, ch) for ch in (8, 12, 16, 24)]
        if k == 4 and sim.fmt == _lib.FMT_F32:
            return [(0, ch) for ch in (8, 12, 14, 16, 20, 24, 32)]
        return []
    out = []
    for strip in (STRIPS if sim.fmt == _lib.FMT_F32 else (128,)):
        if strip - 2 * (k + 1) < 16:
            continue
        for ch in CHUNKS:
            if ch > 2 * H:
                continue
            out.append((strip, min(ch, H)))
    return [(None, None)] + sorted(set(out))


def tune_fused(sim, ks=KS, use_cache=True, reps=10, verbose=False):
    """Best (k, strip, chunk, us_per_step) over ks for `sim`'s lattice; returns (best, table)."""
    cache = load_cache() if use_cache else {}
    bufs = (sim.f.clone(), torch.empty_like(sim.f))
    table = []
    for k in ks:
        key = _key(sim, k)
        if key in cache:
            strip, chunk, us = cache[key]
            table.append((k, strip, chunk, us))
            continue
        best = None
        cands = candidates(sim, k)
        if not cands:
            continue
        for cfg in cands:
            us = time_config(sim, k, cfg, reps=reps, bufs=bufs)
            if verbose:
                print(f"  k={k} strip={cfg[0]} chunk={cfg[1]}: {us:.2f} us/step", flush=True)
            if best is None or us < best[2]:
                best = (cfg[0], cfg[1], us)
        cache[key] = list(best)
        table.append((k,) + tuple(best))
    if use_cache:
        save_cache(cache)
    if not table:
        raise ValueError(f"no fused launch shape for k in {tuple(ks)} from parity {sim.parity}")
    best = min(table, key=lambda r: r[3])
    del bufs
    return best, table
