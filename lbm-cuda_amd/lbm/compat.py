"""Reference-shaped kernel handles over liblbm's C-ABI.

The reference drivers compile a .cu file with CuPy and launch kernels by name with a positional argument
tuple (lbm_smemaa_main.py:84-86,113-123,141-142; lbm_mine2.py:69-80; lbm_smemaa_endfused_main.py:55-58,
72-73,98-101).  ``RawModule(<reference file name>).get_function(<kernel name>)`` returns a callable with the
same ``(grid, block, args)`` signature and the same positional arguments, where arrays are torch tensors
on the GPU instead of CuPy arrays.  ``grid`` and ``block`` are accepted for compatibility and ignored:
launch shapes are chosen by liblbm for gfx950.

Kernel semantics per module (all keep the reference's arguments and buffer roles):
  lbm_core4.cu                 fused_lbmkernel(mask, f_now, f_out, ux, uy, rho, W, H, tau_inv, u_in,
                               tau_inv_bnd): one MRT+Smagorinsky pull step, Cs = 0.2 (lbm_core4.cu:123)
  lbm_core_smemaa.cu           fused_lbmkernel(mask, f_now, f_out, ux, uy, rho, W, H, tau_inv, u_in):
                               T = 4 BGK steps in one launch, then rho/u of the result (lbm_core_smemaa.cu:8)
  lbm_core_smemaa_endfused.cu  fused_lbmkernel(mask, f_now, f_out, W, H, tau_inv, u_in): T = 6 BGK steps
  lbm_core_smemaa_policy.cu    same as endfused (lbm_core_smemaa_policy.cu:7)
  all                          macro_kernel(f, ux, uy, rho, mask, W, H, u_in) (endfused/policy),
                               visualizekernel(ux, uy, rho, image, mask, W, H, vort_scale)
The inlet speed u_in is expanded to the reference's cosine ramp per row (lbm_core4.cu:46-51) on the host
in FP64 and cached on the device.  Deviations from the reference are listed in the project's design notes
(corner cells are walls, the AA kernels' outlet uses lbm_core4's zero-gradient copy).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import geometry as geo

_MODULES = {
    "lbm_core4.cu": dict(kind="pull", model=_lib.MRT_LES, cs=0.2, T=1),
    "lbm_core_smemaa.cu": dict(kind="fused_macros", model=_lib.BGK, cs=0.0, T=4),
    "lbm_core_smemaa_endfused.cu": dict(kind="fused", model=_lib.BGK, cs=0.0, T=6),
    "lbm_core_smemaa_policy.cu": dict(kind="fused", model=_lib.BGK, cs=0.0, T=6),
}

_profiles = {}


def _vp(t):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise TypeError("array arguments must be contiguous torch tensors on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def inlet_profile(H, u_in, device):
    key = (int(H), float(np.float32(u_in)), str(device))
    if key not in _profiles:
        prof = geo.cosine_ramp_profile(int(H), float(np.float32(u_in))).astype(np.float32)
        _profiles[key] = torch.from_numpy(prof).to(device)
    return _profiles[key]


class Kernel:
    def __init__(self, module, name, spec):
        self.module, self.name, self.spec = module, name, spec
        self.lib = _lib.load()

    def __repr__(self):
        return f"<lbm kernel {self.module}:{self.name}>"

    def __call__(self, grid, block, args):
        fn = getattr(self, "_" + self.name)
        return fn(tuple(args))

    def _params(self, W, H, tau_inv):
        sp = self.spec
        return _lib.params(int(W), int(H), float(tau_inv), sp["cs"], sp["model"], _lib.CHANNEL, clamp=1)

    def _fused_lbmkernel(self, a):
        sp = self.spec
        L = self.lib
        if sp["kind"] == "pull":
            mask, f_now, f_out, ux, uy, rho, W, H, tau_inv, u_in, _tau_inv_bnd = a
            p = self._params(W, H, tau_inv)
            prof = inlet_profile(H, u_in, f_now.device)
            _lib.check(L.lbm_step_pull(ctypes.byref(p), _vp(mask), _vp(f_now), _vp(f_out), _vp(ux), _vp(uy),
                                       _vp(rho), _vp(prof), _stream()), "fused_lbmkernel")
            return
        if sp["kind"] == "fused_macros":
            mask, f_now, f_out, ux, uy, rho, W, H, tau_inv, u_in = a
        else:
            mask, f_now, f_out, W, H, tau_inv, u_in = a
        p = self._params(W, H, tau_inv)
        prof = inlet_profile(H, u_in, f_now.device)
        _lib.check(L.lbm_step_fused(ctypes.byref(p), _vp(mask), _vp(f_now), _vp(f_out), _vp(prof), 0, sp["T"], None,
                                    None, None, _stream()), "fused_lbmkernel")
        if sp["kind"] == "fused_macros":
            _lib.check(L.lbm_macro(ctypes.byref(p), _vp(mask), _vp(f_out), _vp(prof), 0, _vp(ux), _vp(uy), _vp(rho),
                                   _stream()), "fused_lbmkernel macros")

    def _macro_kernel(self, a):
        f, ux, uy, rho, mask, W, H, u_in = a
        p = self._params(W, H, 1.0)
        prof = inlet_profile(H, u_in, f.device)
        _lib.check(self.lib.lbm_macro(ctypes.byref(p), _vp(mask), _vp(f), _vp(prof), 0, _vp(ux), _vp(uy), _vp(rho),
                                      _stream()), "macro_kernel")

    def _visualizekernel(self, a):
        ux, uy, _rho, image, mask, W, H, vort_scale = a
        p = self._params(W, H, 1.0)
        _lib.check(self.lib.lbm_visualize(ctypes.byref(p), _vp(ux), _vp(uy), _vp(mask), _vp(image),
                                          float(vort_scale), _stream()), "visualizekernel")


class RawModule:
    """Stand-in for cp.RawModule(code=<reference .cu text>): select the module by its file name."""

    def __init__(self, name):
        name = name.rsplit("/", 1)[-1]
        if name not in _MODULES:
            raise ValueError(f"unknown module {name!r}; known: {sorted(_MODULES)}")
        self.name = name
        self.spec = _MODULES[name]

    def get_function(self, kernel):
        names = {"fused_lbmkernel", "visualizekernel"}
        if self.spec["kind"] == "fused":
            names.add("macro_kernel")
        if kernel not in names:
            raise AttributeError(f"{self.name} has no kernel {kernel!r}")
        return Kernel(self.name, kernel, self.spec)
