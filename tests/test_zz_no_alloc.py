"""GPU: lbm_diagnostics_noalloc (partial sums in the library's static device scratch, no workspace argument)
equals lbm_diagnostics (caller's workspace) bit for bit, through the C-ABI, on two calls in a row -- the second
on a smaller lattice, whose reduction leaves most of the scratch's previous partials in place unless every
workgroup overwrites its own -- and again after a call on another lattice."""
import ctypes

import numpy as np
import pytest
import torch

from lbm import _lib
from lbm import geometry as G

pytestmark = pytest.mark.gpu


def _state(W, H, seed, periodic):
    if periodic:
        mask = G.periodic_mask(W, H)
        mask[H // 3:H // 3 + 4, W // 2:W // 2 + 5] = 1
        ux, uy = G.random_phase_velocity(W, H, 4, 0.04, seed=seed)
        f = G.feq(np.ones((H, W)), ux, uy)
        prof = np.zeros(H)
    else:
        mask = G.channel_mask(W, H, cylinder=(W // 4, H // 2, max(2, H // 8)))
        prof = G.parabolic_profile(H, 0.05)
        f = G.channel_initial_state(W, H, prof, 1e-3, seed)
    dev = torch.device("cuda")
    return (torch.from_numpy(mask).to(dev), torch.from_numpy(f.astype(np.float32)).to(dev),
            torch.from_numpy(prof.astype(np.float32)).to(dev))


def _both(L, W, H, seed, periodic, parity):
    p = _lib.params(W, H, 1.7, cs=0.1, bc=_lib.PERIODIC if periodic else _lib.CHANNEL)
    mask, f, prof = _state(W, H, seed, periodic)
    vp = ctypes.c_void_p
    st = vp(torch.cuda.current_stream().cuda_stream)
    work = torch.empty(L.lbm_diagnostics_work_size(ctypes.byref(p)) // 8, dtype=torch.float64, device="cuda")
    ref = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
    got = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(L.lbm_diagnostics(ctypes.byref(p), vp(mask.data_ptr()), vp(f.data_ptr()), vp(prof.data_ptr()), parity,
                                 vp(work.data_ptr()), vp(ref.data_ptr()), st), "lbm_diagnostics")
    _lib.check(L.lbm_diagnostics_noalloc(ctypes.byref(p), vp(mask.data_ptr()), vp(f.data_ptr()), vp(prof.data_ptr()),
                                         parity, vp(got.data_ptr()), st), "lbm_diagnostics_noalloc")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_diagnostics_equal_workspace_version(gpu):
    L = _lib.load()
    assert L.lbm_diagnostics_scratch_bytes() >= L.lbm_diagnostics_work_size(ctypes.byref(_lib.params(64, 32, 1.7)))
    # a large lattice first (every workgroup of the fixed grid accumulates), then a smaller one, then another
    for W, H, seed, periodic, parity in ((1500, 700, 1, False, 0), (96, 40, 2, False, 1), (300, 120, 3, True, 0),
                                         (64, 16, 4, True, 1)):
        ref, got = _both(L, W, H, seed, periodic, parity)
        assert np.all(np.isfinite(ref)), (W, H)
        assert np.array_equal(ref, got), f"{W}x{H}: {ref} vs {got}"


def test_noalloc_diagnostics_rejects_bad_arguments(gpu):
    L = _lib.load()
    big = _lib.params(40000, 40000, 1.7)
    assert L.lbm_diagnostics_noalloc(ctypes.byref(big), None, None, None, 0, None, None) == -2
    p = _lib.params(64, 32, 1.7)
    assert L.lbm_diagnostics_noalloc(ctypes.byref(p), None, None, None, 0, None, None) == -1
