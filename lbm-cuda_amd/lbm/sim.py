"""Device-resident D2Q9 simulation driven through liblbm's C-ABI.

Mirrors the flow of the reference drivers (lbm_smemaa_main.py:21-174): build the initial state and the
mask on the host, move them to the GPU, then advance with single steps or k-step fused launches.  The
state lives in torch tensors only for device memory and streams; all compute is liblbm's HIP kernels.

Storage: ``f`` is float32 [9, H, W] (SoA planes, x fastest) for ``storage="f32"``; for "f16"/"bf16" it
holds the 16-bit deviations h = f - w (int16 tensor of raw bits), or, with ``row_base``, the deviations from a
per-row base state (lbm_row_base in include/lbm.h).  ``parity`` is the AA parity of the state (0 = canonical,
1 = reversed).  The fused kernel needs a second buffer, allocated on first use.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import geometry as geo

_MODELS = {"mrt": _lib.MRT_LES, "mrt_les": _lib.MRT_LES, "bgk": _lib.BGK}
_BCS = {"channel": _lib.CHANNEL, "periodic": _lib.PERIODIC}
_FMTS = {"f32": _lib.FMT_F32, "f16": _lib.FMT_F16, "bf16": _lib.FMT_BF16}
_CX, _CY, _OPP = geo.CX.astype(np.float64), geo.CY.astype(np.float64), geo.OPP


class LbmHealthError(RuntimeError):
    """A state that failed check_health (NaN/inf, negative density, mass drift)."""


def _ptr(t):
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Simulation:
    """One D2Q9 lattice on one GPU.

    Parameters mirror the reference driver surface: lattice size (W, H), ``tau_inv``, Smagorinsky ``cs``,
    the solid ``mask`` [H, W], the inlet speed (``inlet_u`` per row), steps-per-launch ``k`` (``step``), and
    the AA parity of the stored state.

    ``row_base`` (16-bit storage only; ignored for FP32): store deviations from the equilibrium of a velocity per
    row instead of from the rest state -- "inlet" (the inlet profile), an array ux[H], a pair (ux[H], uy[H]), or
    a ready table [H, 16] (lbm_row_base's layout).  The blocked form is then the two-step kernel (k = 2, even
    parity); single steps take either parity.

    ``window`` = (x0, y0, w, h) (with ``row_base``): inside that rectangle the base is a per-cell one, the
    equilibrium of each cell's own velocity, re-derived from the state every ``rebase_every`` steps (0: only by
    ``rebase()``) -- for the region where the flow departs most from the rows' base (a near wake).  The window lies
    in 1 <= x <= W - 3, 1 <= y <= H - 2; ``rebase_every`` must be a multiple of the k of the launches used.
    ``window="wake"``: the near wake of the channel's bodies (geometry.obstacle_wake_window of the mask).

    ``arith="fast"`` (MRT only): the collision uses the hardware's approximate reciprocal and square root (within 1
    ulp), as the reference's -use_fast_math build does; single steps and the tile kernels (k = 2, 4 from even
    parity) take it, each blocked form bit-identical to its single steps.
    """

    def __init__(self, W, H, tau_inv, *, cs=0.0, model="mrt", bc="channel", mask=None, inlet_u=None, f0=None,
                 clamp=True, s_e=1.1, s_eps=1.1, s_q=1.1, parity=0, device="cuda", storage="f32", sums="plain",
                 row_base=None, arith="exact", window=None, rebase_every=200):
        self.lib = _lib.load()
        self.W, self.H = int(W), int(H)
        self.model = _MODELS[model] if isinstance(model, str) else int(model)
        self.bc = _BCS[bc] if isinstance(bc, str) else int(bc)
        # sums="neumaier": compensated sums of rho - 1 and j in the MRT collision (every kernel)
        # arith="fast": the hardware's approximate reciprocal and square root in the collision (the reference's
        # -use_fast_math arithmetic); "exact": correctly rounded, reproduced bit for bit by the CPU mirror
        self.params = _lib.params(self.W, self.H, tau_inv, cs, self.model, self.bc, int(bool(clamp)), s_e, s_eps,
                                  s_q, {"plain": 0, "neumaier": 1}[sums], fast={"exact": 0, "fast": 1}[arith])
        self.device = torch.device(device)
        if mask is None:
            mask = geo.channel_mask(W, H) if self.bc == _lib.CHANNEL else geo.periodic_mask(W, H)
        elif self.bc == _lib.CHANNEL:
            mask = geo.finalize_channel_mask(mask)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        if mask.shape != (self.H, self.W):
            raise ValueError(f"mask shape {mask.shape} != ({self.H}, {self.W})")
        self.mask_host = mask
        self.mask = torch.from_numpy(mask).to(self.device)
        if inlet_u is None:
            inlet_u = np.zeros(self.H)
        inlet_u = np.asarray(inlet_u, np.float64)
        if inlet_u.shape != (self.H,):
            raise ValueError("inlet_u must have H entries")
        self.inlet_u_host = inlet_u.astype(np.float32)
        self.inlet_u = torch.from_numpy(self.inlet_u_host).to(self.device)
        if f0 is None:
            f0 = geo.feq(np.ones((self.H, self.W)), 0.0, 0.0)
        f0 = np.asarray(f0)
        if f0.shape != (9, self.H, self.W):
            raise ValueError(f"f0 shape {f0.shape} != (9, {self.H}, {self.W})")
        self.fmt = _FMTS[storage] if isinstance(storage, str) else int(storage)
        self.parity = int(parity)
        self.row_base_host = None
        self._row_base = None
        if row_base is not None and self.fmt != _lib.FMT_F32:
            self.row_base_host = self._base_table(row_base)
            self._row_base = torch.from_numpy(self.row_base_host).to(self.device)  # 256-byte aligned
            self.params.row_base = self._row_base.data_ptr()
        self.window = None
        self.rebase_every = int(rebase_every)
        self._since_rebase = 0
        if isinstance(window, str):
            if window != "wake":
                raise ValueError(f"window {window!r}: expected 'wake' or (x0, y0, w, h)")
            window = geo.obstacle_wake_window(mask)
            if window is None:
                raise ValueError("window='wake': the mask holds no body off the wall rows")
        if window is not None:
            if self.row_base_host is None:
                raise ValueError("a base window needs 16-bit storage with row_base")
            x0, y0, w, h = (int(v) for v in window)
            if not (1 <= x0 and x0 + w <= self.W - 2 and 1 <= y0 and y0 + h <= self.H - 1 and w > 0 and h > 0):
                raise ValueError(f"window {window} must lie in 1 <= x <= W - 3, 1 <= y <= H - 2")
            self.window = (x0, y0, w, h)
            # two velocity tables [h, w, 2] of IEEE halves (lbm_rebase_window reads one and writes the other); the
            # first holds each window cell's velocity in f0
            fw = np.asarray(f0, np.float64)[:, y0:y0 + h, x0:x0 + w]
            fw = fw[_OPP] if self.parity else fw
            rho = fw.sum(0)
            vx = (fw * _CX[:, None, None]).sum(0) / rho
            vy = (fw * _CY[:, None, None]).sum(0) / rho
            ok = (rho > 0.5) & (rho < 2.0) & (np.abs(vx) < 1.0) & (np.abs(vy) < 1.0)
            vel = np.stack([np.where(ok, vx, 0.0), np.where(ok, vy, 0.0)], -1).astype(np.float16)
            self._wvel = [torch.from_numpy(vel).to(self.device), torch.empty((h, w, 2), dtype=torch.float16,
                                                                              device=self.device)]
            self._win = _lib.LbmBaseWindow(x0, y0, w, h, self._wvel[0].data_ptr())
            self.params.window = ctypes.pointer(self._win)
        f32 = torch.from_numpy(np.ascontiguousarray(f0, dtype=np.float32)).to(self.device)
        if self.fmt == _lib.FMT_F32:
            self.f = f32
        else:
            # deviations h = f - w (minus the row base) rounded to nearest even on the device
            self.f = torch.empty(f32.shape, dtype=torch.int16, device=self.device)
            _lib.check(self.lib.lbm_convert_state(ctypes.byref(self.params), _ptr(f32), _lib.FMT_F32, _ptr(self.f),
                                                  self.fmt, self.parity, _stream()), "lbm_convert_state")
            del f32
        self.f_alt = None
        self.steps = 0
        self.rho = torch.zeros((self.H, self.W), dtype=torch.float32, device=self.device)
        self.ux = torch.zeros_like(self.rho)
        self.uy = torch.zeros_like(self.rho)
        self._diag_work = None
        self._diag_out = None

    def _base_table(self, row_base):
        if isinstance(row_base, str):
            if row_base != "inlet":
                raise ValueError(f"row_base {row_base!r}: expected 'inlet', ux[H], (ux, uy) or a table [H, 16]")
            return _lib.row_base_table(self.inlet_u_host)
        if isinstance(row_base, tuple):
            ux, uy = row_base
            return _lib.row_base_table(np.asarray(ux, np.float32).reshape(self.H),
                                       np.asarray(uy, np.float32).reshape(self.H))
        a = np.asarray(row_base, np.float32)
        if a.shape == (self.H, _lib.ROW_BASE_STRIDE):
            return np.ascontiguousarray(a)
        if a.shape == (self.H,):
            return _lib.row_base_table(a)
        raise ValueError(f"row_base shape {a.shape}: expected ({self.H},) or ({self.H}, {_lib.ROW_BASE_STRIDE})")

    def window_vel_host(self):
        """The base window's current per-cell base velocities [h, w, 2] (float16; None without a window)."""
        return None if self.window is None else self._wvel[0].cpu().numpy()

    def rebase(self):
        """Re-derive the base window's per-cell base velocities from the stored state (lbm_rebase_window): the
        stored words are re-encoded against the equilibrium of each window cell's current velocity."""
        if self.window is None:
            return self
        rc = self.lib.lbm_rebase_window(ctypes.byref(self.params), _ptr(self.f), self.fmt, self.parity,
                                        _ptr(self._wvel[1]), _stream())
        _lib.check(rc, "lbm_rebase_window")
        self._wvel.reverse()
        self._win.vel = self._wvel[0].data_ptr()
        self._since_rebase = 0
        return self

    # ------------------------------------------------------------------ stepping
    def step(self, n=1, k=1, macros=False, cfg=None):
        """Advance n collide-stream steps; k > 1 uses the temporally blocked kernel (k steps per launch).
        ``cfg`` = (strip, chunk) selects its launch configuration (default, or (None, None) as the tuner
        returns it: the library's choice).  Runs of
        several launches are issued by one C call (lbm_step_aa_n / lbm_step_fused_n); with ``macros`` only
        the last launch writes them.  With a base window, the run is cut at every ``rebase_every`` steps for a
        re-derivation of the window's base."""
        n = int(n)
        k = int(k)
        if self.window is not None and self.rebase_every > 0:
            if self.rebase_every % max(k, 1) or self._since_rebase % max(k, 1):
                raise ValueError(f"rebase_every={self.rebase_every} (at {self._since_rebase} since the last) is "
                                 f"not a multiple of k={k}")
            while n > 0:
                m = min(n, self.rebase_every - self._since_rebase)
                self._step(m, k, macros and m == n, cfg)
                self._since_rebase += m
                n -= m
                if self._since_rebase >= self.rebase_every:
                    self.rebase()
            return self
        return self._step(n, k, macros, cfg)

    def _step(self, n, k, macros, cfg):
        if k <= 1:
            if n > 1:
                self._aa_n(n - 1 if macros else n)
                if macros:
                    self._aa(True)
            elif n == 1:
                self._aa(macros)
            return self
        if n % k:
            raise ValueError(f"n={n} is not a multiple of k={k}")
        nl = n // k
        if nl > 1:
            self._fused_n(k, nl - 1 if macros else nl, cfg)
            if macros:
                self._fused(k, True, cfg)
        elif nl == 1:
            self._fused(k, macros, cfg)
        return self

    def _aa_n(self, n):
        if self.fmt == _lib.FMT_F32:
            rc = self.lib.lbm_step_aa_n(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), _ptr(self.inlet_u),
                                        self.parity, n, _stream())
        else:
            rc = self.lib.lbm_step_aa_h16_n(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), self.fmt,
                                            _ptr(self.inlet_u), self.parity, n, _stream())
        _lib.check(rc, "lbm_step_aa_n")
        self.parity = (self.parity + n) & 1
        self.steps += n

    def _fused_n(self, k, nl, cfg=None):
        if self.f_alt is None:
            self.f_alt = torch.empty_like(self.f)
        c = None if cfg is None or cfg[0] is None else ctypes.byref(_lib.LbmFusedCfg(int(cfg[0]), int(cfg[1])))
        if self.fmt == _lib.FMT_F32:
            rc = self.lib.lbm_step_fused_n(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), _ptr(self.f_alt),
                                           _ptr(self.inlet_u), self.parity, k, nl, c, _stream())
        else:
            rc = self.lib.lbm_step_fused_h16_n(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f),
                                               _ptr(self.f_alt), self.fmt, _ptr(self.inlet_u), self.parity, k, nl, c,
                                               _stream())
        _lib.check(rc, "lbm_step_fused_n")
        if nl & 1:
            self.f, self.f_alt = self.f_alt, self.f
        self.parity = (self.parity + k * nl) & 1
        self.steps += k * nl

    def _aa(self, macros=False):
        m = (self.ux, self.uy, self.rho) if macros else (None, None, None)
        if self.fmt == _lib.FMT_F32:
            rc = self.lib.lbm_step_aa(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), _ptr(self.inlet_u),
                                      self.parity, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]), _stream())
        else:
            rc = self.lib.lbm_step_aa_h16(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), self.fmt,
                                          _ptr(self.inlet_u), self.parity, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]),
                                          _stream())
        _lib.check(rc, "lbm_step_aa")
        self.parity ^= 1
        self.steps += 1

    def _fused(self, k, macros=False, cfg=None):
        if self.f_alt is None:
            self.f_alt = torch.empty_like(self.f)
        m = (self.ux, self.uy, self.rho) if macros else (None, None, None)
        c = None if cfg is None or cfg[0] is None else ctypes.byref(_lib.LbmFusedCfg(int(cfg[0]), int(cfg[1])))
        if self.fmt == _lib.FMT_F32:
            rc = self.lib.lbm_step_fused_cfg(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f),
                                             _ptr(self.f_alt), _ptr(self.inlet_u), self.parity, k, c, _ptr(m[0]),
                                             _ptr(m[1]), _ptr(m[2]), _stream())
        else:
            rc = self.lib.lbm_step_fused_h16(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f),
                                             _ptr(self.f_alt), self.fmt, _ptr(self.inlet_u), self.parity, k, c,
                                             _ptr(m[0]), _ptr(m[1]), _ptr(m[2]), _stream())
        _lib.check(rc, "lbm_step_fused")
        self.f, self.f_alt = self.f_alt, self.f
        self.parity = (self.parity + k) & 1
        self.steps += k

    # ------------------------------------------------------------------ observables
    def macro(self):
        """(rho, ux, uy) of the stored state (torch tensors on the device)."""
        if self.fmt == _lib.FMT_F32:
            rc = self.lib.lbm_macro(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), _ptr(self.inlet_u),
                                    self.parity, _ptr(self.ux), _ptr(self.uy), _ptr(self.rho), _stream())
        else:
            rc = self.lib.lbm_macro_h16(ctypes.byref(self.params), _ptr(self.mask), _ptr(self.f), self.fmt,
                                        _ptr(self.inlet_u), self.parity, _ptr(self.ux), _ptr(self.uy), _ptr(self.rho),
                                        _stream())
        _lib.check(rc, "lbm_macro")
        return self.rho, self.ux, self.uy

    def populations(self):
        """FP32 populations f of the stored state (a device tensor; converted for 16-bit storage)."""
        if self.fmt == _lib.FMT_F32:
            return self.f
        out = torch.empty(self.f.shape, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.lbm_convert_state(ctypes.byref(self.params), _ptr(self.f), self.fmt, _ptr(out),
                                              _lib.FMT_F32, self.parity, _stream()), "lbm_convert_state")
        return out

    def diagnostics(self):
        """FP64 (mass, jx, jy, Fx, Fy) reduced on the GPU."""
        if self._diag_work is None:
            nbytes = self.lib.lbm_diagnostics_work_size(ctypes.byref(self.params))
            self._diag_work = torch.empty(nbytes // 8, dtype=torch.float64, device=self.device)
            self._diag_out = torch.empty(5, dtype=torch.float64, device=self.device)
        f = self.populations()
        rc = self.lib.lbm_diagnostics(ctypes.byref(self.params), _ptr(self.mask), _ptr(f), _ptr(self.inlet_u),
                                      self.parity, _ptr(self._diag_work), _ptr(self._diag_out), _stream())
        _lib.check(rc, "lbm_diagnostics")
        return self._diag_out.cpu().numpy().copy()

    def check_health(self, mass_ref=None, rel_tol=1e-3):
        """Failure detection: raise LbmHealthError when the state holds a non-finite value, a negative density,
        or (given mass_ref) a mass that drifted by more than rel_tol; otherwise return the figures checked.
        Mass and momentum come from the FP64 reduction on the GPU, the density minimum from lbm_macro."""
        d = self.diagnostics()
        if not np.all(np.isfinite(d)):
            raise LbmHealthError(f"non-finite state after {self.steps} steps (mass {d[0]}, momentum {d[1:3]})")
        rho = self.macro()[0].cpu().numpy()
        upd = self.mask_host == 0
        if self.bc == _lib.CHANNEL:
            upd = upd.copy()
            upd[0, :] = upd[-1, :] = False
            upd[:, -1] = False
        rho_min = float(rho[upd].min()) if upd.any() else 1.0
        if not np.isfinite(rho_min) or rho_min < 0.0:
            raise LbmHealthError(f"negative or non-finite density {rho_min} after {self.steps} steps")
        drift = None
        if mass_ref is not None:
            drift = abs(d[0] - mass_ref) / abs(mass_ref)
            if drift > rel_tol:
                raise LbmHealthError(f"mass drifted by {drift:.3e} (> {rel_tol}) after {self.steps} steps")
        return {"mass": float(d[0]), "momentum": (float(d[1]), float(d[2])), "rho_min": rho_min, "mass_drift": drift}

    def drag_lift(self, u_ref, length, rho_ref=1.0):
        """(Cd, Cl) = 2 F / (rho_ref u_ref^2 length) from the momentum-exchange force."""
        d = self.diagnostics()
        scale = 2.0 / (rho_ref * u_ref * u_ref * length)
        return d[3] * scale, d[4] * scale

    def visualize(self, vort_scale=50.0):
        """RGBA8 vorticity image [H, W, 4] (interior cells; computed from the current macro fields)."""
        self.macro()
        img = torch.zeros((self.H, self.W, 4), dtype=torch.uint8, device=self.device)
        rc = self.lib.lbm_visualize(ctypes.byref(self.params), _ptr(self.ux), _ptr(self.uy), _ptr(self.mask),
                                    _ptr(img), float(vort_scale), _stream())
        _lib.check(rc, "lbm_visualize")
        return img

    def state(self):
        """(stored state as numpy [9, H, W] -- float32 f, or uint16 deviation bits -- , parity)."""
        a = self.f.cpu().numpy()
        return (a if self.fmt == _lib.FMT_F32 else a.view(np.uint16)), self.parity

    def synchronize(self):
        _lib.check(self.lib.lbm_sync(_stream()), "lbm_sync")

    # ------------------------------------------------------------------ checkpoint / resume
    def save(self, path):
        """Write f, parity, step count and parameters to an .npz file."""
        p = self.params
        extra = {} if self.row_base_host is None else {"row_base": self.row_base_host}
        if self.window is not None:
            extra.update(window=np.array(self.window, np.int64), window_vel=self.window_vel_host(),
                         rebase=np.array([self.rebase_every, self._since_rebase], np.int64))
        np.savez(path, f=self.populations().cpu().numpy(), parity=self.parity, steps=self.steps, mask=self.mask_host,
                 inlet_u=self.inlet_u_host, raw=self.f.cpu().numpy(),
                 params=np.array([p.W, p.H, p.model, p.bc, p.clamp, self.fmt], np.int64),
                 fparams=np.array([p.tau_inv, p.cs, p.s_e, p.s_eps, p.s_q], np.float32), fast=p.fast, **extra)

    @classmethod
    def load(cls, path, device="cuda"):
        z = np.load(path, allow_pickle=False)
        W, H, model, bc, clamp, fmt = (int(v) for v in z["params"])
        tau_inv, cs, s_e, s_eps, s_q = (float(v) for v in z["fparams"])
        sim = cls(W, H, tau_inv, cs=cs, model=model, bc=bc, mask=z["mask"], inlet_u=z["inlet_u"], f0=z["f"],
                  clamp=bool(clamp), s_e=s_e, s_eps=s_eps, s_q=s_q, parity=int(z["parity"]), device=device,
                  storage=fmt, row_base=z["row_base"] if "row_base" in z.files else None,
                  arith="fast" if ("fast" in z.files and int(z["fast"])) else "exact",
                  window=z["window"] if "window" in z.files else None,
                  rebase_every=int(z["rebase"][0]) if "rebase" in z.files else 200)
        sim.f.copy_(torch.from_numpy(z["raw"]).to(sim.device))  # exact stored bits
        if sim.window is not None:
            sim._wvel[0].copy_(torch.from_numpy(z["window_vel"]).to(sim.device))
            sim._since_rebase = int(z["rebase"][1])
        sim.steps = int(z["steps"])
        return sim
