"""Learned subgrid closure: an MLP 18 -> 64 -> 64 -> 1 (GELU) predicting the eddy viscosity nu_t of a coarse
lattice from the 3 x 3 stencil of its non-equilibrium stress moments (m7neq, m8neq), trained on a filtered
fine run and evaluated inside the collide-stream step, with Smagorinsky as the fallback.

The reference's only closure is the Smagorinsky model of lbm_core4.cu:171-177, driven by the same two moments
(lbm_core4.cu:172-173); everything here is new.  All GPU work goes through liblbm's C-ABI:
  lbm_mlp_pack / lbm_mlp_forward / lbm_mlp_grad / lbm_mlp_train_step   BF16 matrix-core kernels, FP32 master
  lbm_neq / lbm_step_closure                                           features and the fused closure step
Host-side numpy does the data preparation and analysis only: the box filter, the subgrid-stress targets,
energy spectra and enstrophy.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .sim import Simulation, _ptr, _stream

NIN, NH = 18, 64
NPARAM = NIN * NH + NH + NH * NH + NH + NH + 1  # 5441


def init_params(seed=0):
    """He-style initial parameters (FP32), in the kernels' order W1, b1, W2, b2, W3, b3."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    p[o:o + NIN * NH] = rng.normal(0, math.sqrt(2.0 / NIN), NIN * NH)
    o += NIN * NH + NH
    p[o:o + NH * NH] = rng.normal(0, math.sqrt(2.0 / NH), NH * NH)
    o += NH * NH + NH
    p[o:o + NH] = rng.normal(0, math.sqrt(1.0 / NH), NH)
    return p


class MLP:
    """The network's FP32 parameters, packed BF16 weights and Adam state, resident on the GPU."""

    def __init__(self, params=None, device="cuda", seed=0):
        self.lib = _lib.load()
        if self.lib.lbm_mlp_nparam() != NPARAM:
            raise _lib.LbmError("liblbm's MLP parameter count differs")
        self.device = torch.device(device)
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = torch.from_numpy(p.copy()).to(self.device)
        self.packed = torch.empty(self.lib.lbm_mlp_npack(), dtype=torch.int16, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0
        self._work = None
        self.repack()

    def repack(self):
        _lib.check(self.lib.lbm_mlp_pack(_ptr(self.params), _ptr(self.packed), _stream()), "lbm_mlp_pack")

    def _work_for(self, n):
        need = self.lib.lbm_mlp_work_size(int(n))
        if self._work is None or self._work.numel() * 4 < need:
            self._work = torch.empty(need // 4, dtype=torch.float32, device=self.device)
        return self._work

    def forward(self, x):
        """Network output for samples x [n, 18] (device float32 tensor)."""
        x = x.contiguous()
        y = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        _lib.check(self.lib.lbm_mlp_forward(_ptr(self.packed), _ptr(self.params), _ptr(x), int(x.shape[0]), _ptr(y),
                                            _stream()), "lbm_mlp_forward")
        return y

    def grad(self, x, t, w):
        """(gradient [5441], loss) of sum_s w_s (y_s - t_s)^2."""
        n = int(x.shape[0])
        g = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.lbm_mlp_grad(_ptr(self.packed), _ptr(self.params), _ptr(x.contiguous()),
                                         _ptr(t.contiguous()), _ptr(w.contiguous()), n, _ptr(self._work_for(n)),
                                         _ptr(g), _ptr(loss), _stream()), "lbm_mlp_grad")
        return g, loss

    def train_step(self, x, t, w, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, grad=None, loss=None):
        """One Adam step on the weighted squared error of this batch."""
        n = int(x.shape[0])
        self.t += 1
        hp = _lib.LbmAdam(lr, beta1, beta2, eps, 1.0 - beta1 ** self.t, 1.0 - beta2 ** self.t)
        _lib.check(self.lib.lbm_mlp_train_step(_ptr(self.packed), _ptr(self.params), _ptr(self.m), _ptr(self.v),
                                               _ptr(x.contiguous()), _ptr(t.contiguous()), _ptr(w.contiguous()), n,
                                               ctypes.byref(hp), _ptr(self._work_for(n)), _ptr(grad), _ptr(loss),
                                               _stream()), "lbm_mlp_train_step")
        return hp

    def numpy(self):
        return self.params.cpu().numpy().copy()


@dataclass
class ClosureConfig:
    """Feature scaling and trust region (lbm_closure in include/lbm.h)."""
    in_scale: np.ndarray = field(default_factory=lambda: np.ones(NIN, np.float32))
    nu_scale: float = 1.0
    nu_max: float = 0.05
    feat_max: float = 8.0

    def c_struct(self):
        c = _lib.LbmClosure()
        for k in range(NIN):
            c.in_scale[k] = float(self.in_scale[k])
        c.nu_scale, c.nu_max, c.feat_max = float(self.nu_scale), float(self.nu_max), float(self.feat_max)
        return c


class ClosureSim:
    """A periodic MRT lattice stepped with the learned closure (lbm_step_closure).  ``sim`` supplies the
    state, mask and parameters (its cs is the fallback's Smagorinsky constant)."""

    def __init__(self, sim: Simulation, mlp: MLP, cfg: ClosureConfig):
        if sim.bc != _lib.PERIODIC or sim.fmt != _lib.FMT_F32 or sim.model != _lib.MRT_LES:
            raise ValueError("the learned closure runs on periodic FP32 MRT lattices")
        self.sim, self.mlp, self.cfg = sim, mlp, cfg
        self._c = cfg.c_struct()
        self.neq = torch.empty((2, sim.H, sim.W), dtype=torch.float32, device=sim.device)
        self.nut = torch.empty((sim.H, sim.W), dtype=torch.float32, device=sim.device)
        self.fallbacks = torch.zeros(1, dtype=torch.int32, device=sim.device)

    def step(self, n=1, record_nut=False):
        s = self.sim
        for _ in range(int(n)):
            _lib.check(s.lib.lbm_step_closure(ctypes.byref(s.params), _ptr(s.mask), _ptr(s.f), s.parity,
                                              _ptr(self.mlp.packed), _ptr(self.mlp.params), ctypes.byref(self._c),
                                              _ptr(self.neq), _ptr(self.nut) if record_nut else None,
                                              _ptr(self.fallbacks), _stream()), "lbm_step_closure")
            s.parity ^= 1
            s.steps += 1
        return self


def neq(sim: Simulation):
    """(m7neq, m8neq) of the simulation's current state, [2, H, W] device tensor."""
    out = torch.empty((2, sim.H, sim.W), dtype=torch.float32, device=sim.device)
    _lib.check(sim.lib.lbm_neq(ctypes.byref(sim.params), _ptr(sim.mask), _ptr(sim.f), sim.parity, _ptr(out),
                               _stream()), "lbm_neq")
    return out


def stencil_features(neq_hw):
    """The 18 features of every cell from a [2, H, W] numpy array: feature 2 s + c is moment c at
    (x + dx, y + dy), s = 3 (dy + 1) + (dx + 1), periodic; returns [H*W, 18]."""
    cols = []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            for c in range(2):
                cols.append(np.roll(neq_hw[c], (-dy, -dx), axis=(0, 1)).reshape(-1))
    return np.stack(cols, 1)
