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


# ------------------------------------------------------------------------------------------------ the study
def box_filter(a, factor):
    """Mean over factor x factor blocks (the coarse cell (i, j) covers fine cells i*factor .. i*factor+factor-1)."""
    H, W = a.shape
    return a.reshape(H // factor, factor, W // factor, factor).mean(axis=(1, 3))


def filtered_fields(sim: Simulation, factor):
    """Favre-filtered rho, u, v and the subgrid stress tau_ab = (rho u_a u_b)~ / rho~ - u~_a u~_b of a fine run
    on the coarse grid (FP64, host)."""
    rho, ux, uy = (t.cpu().numpy().astype(np.float64) for t in sim.macro())
    rb = box_filter(rho, factor)
    u = box_filter(rho * ux, factor) / rb
    v = box_filter(rho * uy, factor) / rb
    txx = box_filter(rho * ux * ux, factor) / rb - u * u
    tyy = box_filter(rho * uy * uy, factor) / rb - v * v
    txy = box_filter(rho * ux * uy, factor) / rb - u * v
    return rb, u, v, (txx, tyy, txy)


def strain(u, v):
    """Deviatoric strain (s1 = (S_xx - S_yy) / 2, s2 = S_xy) by periodic central differences (lattice units)."""
    dudx = 0.5 * (np.roll(u, -1, 1) - np.roll(u, 1, 1))
    dudy = 0.5 * (np.roll(u, -1, 0) - np.roll(u, 1, 0))
    dvdx = 0.5 * (np.roll(v, -1, 1) - np.roll(v, 1, 1))
    dvdy = 0.5 * (np.roll(v, -1, 0) - np.roll(v, 1, 0))
    return 0.5 * (dudx - dvdy), 0.5 * (dudy + dvdx)


def window_sum(a, w):
    """Periodic sum over the w x w neighbourhood centred on each cell (w odd)."""
    out = np.zeros_like(a)
    r = w // 2
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            out += np.roll(a, (dy, dx), axis=(0, 1))
    return out


def sgs_target(u, v, tau, window=1):
    """Eddy viscosity that best matches the deviatoric subgrid stress in the least-squares sense over the
    window x window neighbourhood of each cell, nu* = -sum(d : S) / (2 sum(S : S)), and the weight sum(S : S)
    of that fit (sum |tau_d + 2 nu S|^2 = 4 sum|S|^2 (nu - nu*)^2 + c).  window = 1 is the point-wise fit."""
    s1, s2 = strain(u, v)
    txx, tyy, txy = tau
    d1, d2 = 0.5 * (txx - tyy), txy
    num = -(d1 * s1 + d2 * s2)
    ss = s1 * s1 + s2 * s2
    if window > 1:
        num, ss = window_sum(num, window), window_sum(ss, window)
    return num / (2.0 * ss + 1e-30), ss


def smagorinsky_nu(neq_hw, rho, tau0, cs):
    """nu_t = (tau_eff - tau_0) / 3 of the Smagorinsky rule of lbm_core4.cu:174-176 for given moments."""
    Q = np.sqrt(neq_hw[0].astype(np.float64) ** 2 + neq_hw[1].astype(np.float64) ** 2 + 1e-8)
    tau_eff = 0.5 * (tau0 + np.sqrt(tau0 * tau0 + 18.0 * cs * cs * Q / rho))
    return (tau_eff - tau0) / 3.0


def spectrum(u, v):
    """Shell-averaged kinetic energy spectrum E(k), k = 0 .. N/2 (integer shells), of a periodic field."""
    H, W = u.shape
    uk = np.fft.fft2(u) / (H * W)
    vk = np.fft.fft2(v) / (H * W)
    e = 0.5 * (np.abs(uk) ** 2 + np.abs(vk) ** 2)
    kx = np.fft.fftfreq(W) * W
    ky = np.fft.fftfreq(H) * H
    k = np.rint(np.hypot(*np.meshgrid(kx, ky))).astype(int)
    kmax = min(H, W) // 2
    return np.bincount(k.ravel(), weights=e.ravel(), minlength=kmax + 1)[:kmax + 1]


def enstrophy(u, v):
    """Z = <omega^2> / 2 with omega = dv/dx - du/dy (spectral derivatives)."""
    H, W = u.shape
    kx = 2j * np.pi * np.fft.fftfreq(W)[None, :]
    ky = 2j * np.pi * np.fft.fftfreq(H)[:, None]
    w = np.real(np.fft.ifft2(kx * np.fft.fft2(v) - ky * np.fft.fft2(u)))
    return 0.5 * float(np.mean(w * w))


@dataclass
class StudyConfig:
    """The learned-closure workload: fine N x N periodic run (MRT + Smagorinsky cs_fine, random-phase field
    with E(k) ~ k^4 exp(-2 (k / k0)^2) at u_rms), box filter by `factor`, coarse runs with nu / factor
    (acoustic scaling: one coarse step = `factor` fine steps)."""
    N: int = 4096
    factor: int = 8
    nu_fine: float = 1e-4
    cs_fine: float = 0.05
    cs_coarse: float = 0.1     # the Smagorinsky baseline and the closure's fallback
    k0: float = 16.0
    u_rms: float = 0.05
    seed: int = 5
    spinup: int = 2000         # fine steps before the first snapshot
    snapshots: int = 8         # training snapshots (the last one is held out for validation)
    interval: int = 400        # fine steps between snapshots
    warm: int = 10             # coarse Smagorinsky steps before the features are taken
    window: int = 33           # neighbourhood of the least-squares eddy-viscosity target (coarse cells)
    clip_target: bool = False  # train on max(nu*, 0) (the step clamps negative outputs to 0 anyway)
    augment: bool = True       # train on the 8 rotations / reflections of every sample (the D4 symmetry)
    eval_steps: int = 20000    # coarse steps of the a-posteriori comparison
    eval_every: int = 2000
    epochs: int = 6
    batch: int = 8192
    lr: float = 2e-3


def _step_fast(sim, n):
    """n steps of an FP32 periodic lattice, four per launch with the four-step kernel where the parity allows
    (bit-identical to single steps)."""
    n = int(n)
    if sim.parity:
        sim.step(1)
        n -= 1
    q = (n // 4) * 4
    if q:
        sim.step(q, k=4, cfg=(0, 24))
    if n - q:
        sim.step(n - q)


def _coarse_sim(cfg: StudyConfig, rho, u, v, cs, device):
    n = rho.shape[0]
    nu = cfg.nu_fine / cfg.factor
    tau_inv = float(np.float32(1.0 / (3.0 * nu + 0.5)))
    return Simulation(n, n, tau_inv, cs=cs, model="mrt", bc="periodic", f0=regularized_state(rho, u, v, nu, cs),
                      device=device)


def regularized_state(rho, u, v, nu, cs):
    """Populations with the Chapman-Enskog non-equilibrium part of the given velocity field:
    f_q = feq_q - 3 w_q rho tau (c_q c_q - I / 3) : S, tau = 3 (nu + cs^2 |S|) + 1/2, |S| = sqrt(2 S : S).
    Starting a lattice near tau = 1/2 from feq alone leaves a non-equilibrium transient that changes sign every
    step and decays only as |1 - 1/tau|^n."""
    from .geometry import CX, CY, WEIGHTS, feq
    f = feq(rho, u, v)
    dudx = 0.5 * (np.roll(u, -1, 1) - np.roll(u, 1, 1))
    dudy = 0.5 * (np.roll(u, -1, 0) - np.roll(u, 1, 0))
    dvdx = 0.5 * (np.roll(v, -1, 1) - np.roll(v, 1, 1))
    dvdy = 0.5 * (np.roll(v, -1, 0) - np.roll(v, 1, 0))
    sxx, syy, sxy = dudx, dvdy, 0.5 * (dudy + dvdx)
    tau = 3.0 * (nu + cs * cs * np.sqrt(2.0 * (sxx * sxx + syy * syy + 2.0 * sxy * sxy))) + 0.5
    tr = (sxx + syy) / 3.0
    for q in range(9):
        cs_q = CX[q] * CX[q] * sxx + 2.0 * CX[q] * CY[q] * sxy + CY[q] * CY[q] * syy - tr
        f[q] -= 3.0 * WEIGHTS[q] * rho * tau * cs_q
    return f


def _coarse_uv(sim):
    rho, ux, uy = (t.cpu().numpy().astype(np.float64) for t in sim.macro())
    return rho, ux, uy


def d4_augment(X, T, Wt):
    """The 8 images of every sample under the symmetries of the square lattice.  Feature 2 s + c holds moment c
    (c = 0: m7neq ~ S_xx - S_yy, c = 1: m8neq ~ S_xy) at stencil offset s = 3 (dy + 1) + (dx + 1).  A rotation
    by 90 degrees maps the offset (dx, dy) to (-dy, dx) and flips the signs of both moments; the reflection
    x -> -x maps (dx, dy) to (-dx, dy) and flips the sign of m8neq.  The target (a scalar viscosity) and the
    weight are invariant."""
    def perm(fn):
        out = np.empty(18, np.int64)
        for s in range(9):
            dx, dy = s % 3 - 1, s // 3 - 1
            ndx, ndy = fn(dx, dy)
            t = 3 * (ndy + 1) + (ndx + 1)
            out[2 * t], out[2 * t + 1] = 2 * s, 2 * s + 1
        return out
    rot = perm(lambda dx, dy: (-dy, dx))
    ref = perm(lambda dx, dy: (-dx, dy))
    s_rot = np.tile(np.array([-1.0, -1.0], np.float32), 9)
    s_ref = np.tile(np.array([1.0, -1.0], np.float32), 9)
    xs = []
    base = X
    for _ in range(4):
        xs.append(base)
        xs.append(base[:, ref] * s_ref[None, :])
        base = base[:, rot] * s_rot[None, :]
    return np.concatenate(xs), np.tile(T, 8), np.tile(Wt, 8)


def build_dataset(cfg: StudyConfig, device="cuda", log=print):
    """The fine run (spin-up, then `snapshots` snapshots `interval` steps apart) and, per snapshot, the coarse
    features (3 x 3 (m7neq, m8neq) stencils of a coarse lattice started from the filtered state after `warm`
    Smagorinsky steps), the windowed least-squares eddy-viscosity target of the filtered fine subgrid stress,
    its weight, and the Smagorinsky viscosity of the same coarse state.  Returns (fine, dataset dict)."""
    import time
    from . import geometry as geo
    t_start = time.perf_counter()
    N, fct = cfg.N, cfg.factor
    ux, uy = geo.random_phase_velocity(N, N, k0=cfg.k0, u_rms=cfg.u_rms, seed=cfg.seed)
    fine = Simulation(N, N, float(np.float32(geo.nu_tau_inv(cfg.nu_fine))), cs=cfg.cs_fine, model="mrt",
                      bc="periodic", f0=geo.feq(np.ones((N, N)), ux, uy), device=device)
    del ux, uy
    log(f"closure study: fine {N}^2 built in {time.perf_counter() - t_start:.1f} s")
    _step_fast(fine, cfg.spinup)
    X, T, Wt, smag_nu = [], [], [], []
    for s in range(cfg.snapshots):
        rb, u, v, _ = filtered_fields(fine, fct)
        coarse = _coarse_sim(cfg, rb, u, v, cfg.cs_coarse, device)
        coarse.step(cfg.warm)
        nq = neq(coarse).cpu().numpy()
        _step_fast(fine, cfg.warm * fct)
        rb2, u2, v2, tau2 = filtered_fields(fine, fct)
        nu_star, ss = sgs_target(u2, v2, tau2, cfg.window)
        X.append(stencil_features(nq).astype(np.float32))
        T.append(nu_star.reshape(-1))
        Wt.append(ss.reshape(-1))
        rho_c = coarse.macro()[0].cpu().numpy().astype(np.float64)
        smag_nu.append(smagorinsky_nu(nq, rho_c, 1.0 / coarse.params.tau_inv, cfg.cs_coarse).reshape(-1))
        del coarse
        _step_fast(fine, cfg.interval)
    log(f"closure study: {cfg.snapshots} snapshots in {time.perf_counter() - t_start:.1f} s")
    return fine, {"X": X, "T": T, "W": Wt, "smag": smag_nu}


def train_closure(cfg: StudyConfig, data, device="cuda", log=print):
    """Normalise the training snapshots (all but the last), train the MLP with the GPU kernels (Adam, cosine
    learning rate, weighted squared error), and score it a priori on the held-out snapshot.  Returns
    (mlp, closure config, a-priori report, training info)."""
    import time
    X, T, Wt, smag_nu = data["X"], data["T"], data["W"], data["smag"]
    Xtr = np.concatenate(X[:-1])
    s7 = 1.0 / Xtr[:, 0::2].std()
    s8 = 1.0 / Xtr[:, 1::2].std()
    s78 = 1.0 / np.sqrt(0.5 * (Xtr[:, 0::2].var() + Xtr[:, 1::2].var())) if cfg.augment else None
    in_scale = np.tile(np.array([s78, s78] if cfg.augment else [s7, s8], np.float32), 9)
    Ttr = np.concatenate(T[:-1])
    Wtr = np.concatenate(Wt[:-1])
    nu_scale = float(np.sqrt(np.sum(Wtr * Ttr ** 2) / np.sum(Wtr)))
    ccfg = ClosureConfig(in_scale=in_scale, nu_scale=nu_scale, nu_max=0.05, feat_max=20.0)

    def prep(Xs, Ts, Ws, aug):
        x = np.concatenate(Xs)
        t = np.concatenate(Ts)
        w = np.concatenate(Ws)
        if cfg.clip_target:
            t = np.maximum(t, 0.0)
        if aug:
            x, t, w = d4_augment(x, t, w)
        x = x * in_scale[None, :]
        t = t / nu_scale
        return x.astype(np.float32), t.astype(np.float32), (w / w.mean()).astype(np.float32)

    xtr, ttr, wtr = prep(X[:-1], T[:-1], Wt[:-1], cfg.augment)
    xva, tva, wva = prep(X[-1:], T[-1:], Wt[-1:], False)
    tva_raw = np.concatenate(T[-1:]) / nu_scale  # the unclipped target, for the a-priori scores
    mlp = MLP(seed=cfg.seed, device=device)
    y_init = mlp.forward(torch.from_numpy(xva).to(device)).cpu().numpy()
    rng = np.random.default_rng(cfg.seed)
    nb = len(xtr) // cfg.batch
    loss = torch.empty(1, dtype=torch.float32, device=device)
    t_train = time.perf_counter()
    steps = 0
    for ep in range(cfg.epochs):
        perm = rng.permutation(len(xtr))[:nb * cfg.batch]
        xe = torch.from_numpy(xtr[perm]).to(device)
        te = torch.from_numpy(ttr[perm]).to(device)
        we = torch.from_numpy(wtr[perm] / cfg.batch).to(device)
        lr = cfg.lr * 0.5 * (1 + math.cos(math.pi * ep / cfg.epochs))
        for b in range(nb):
            sl = slice(b * cfg.batch, (b + 1) * cfg.batch)
            mlp.train_step(xe[sl], te[sl], we[sl], lr=lr, loss=loss)
            steps += 1
        log(f"  epoch {ep}: lr {lr:.2e}, last batch loss {loss.item():.4f}")
        del xe, te, we
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t_train

    def wmse(pred, t, w):
        return float(np.sum(w * (pred - t) ** 2) / np.sum(w))

    y_va = mlp.forward(torch.from_numpy(xva).to(device)).cpu().numpy()
    smag_va = smag_nu[-1] / nu_scale
    y_used = np.maximum(y_va, 0.0)  # what the step applies
    apriori = {"val_wmse_mlp": wmse(y_va, tva_raw, wva), "val_wmse_mlp_clamped": wmse(y_used, tva_raw, wva),
               "val_wmse_untrained": wmse(y_init, tva_raw, wva),
               "val_wmse_smagorinsky": wmse(smag_va, tva_raw, wva),
               "val_wmse_zero": wmse(np.zeros_like(tva_raw), tva_raw, wva),
               "corr_mlp": float(np.corrcoef(y_va, tva_raw)[0, 1]),
               "corr_smagorinsky": float(np.corrcoef(smag_va, tva_raw)[0, 1]),
               "mean_nu_ratio_mlp": float(np.sum(wva * y_used) / np.sum(wva * np.maximum(tva_raw, 0.0))),
               "mean_nu_ratio_smagorinsky": float(np.sum(wva * smag_va) / np.sum(wva * np.maximum(tva_raw, 0.0)))}
    log(f"closure study: trained {steps} steps in {train_s:.1f} s; a-priori {apriori}")
    info = {"samples_train": int(len(xtr)), "samples_val": int(len(xva)), "train_steps": steps,
            "train_seconds": round(train_s, 2), "nu_scale": nu_scale}
    return mlp, ccfg, apriori, info


def reference_track(cfg: StudyConfig, fine: Simulation, log=print):
    """The filtered fine run from now on: its filtered start state (rho, u, v) for the coarse runs and, every
    `eval_every` coarse steps up to `eval_steps`, the filtered field's energy spectrum, energy and enstrophy."""
    fct = cfg.factor
    rb, u, v, _ = filtered_fields(fine, fct)
    start = (rb, u, v)
    track = []
    for k in range(0, cfg.eval_steps + 1, cfg.eval_every):
        if k:
            _step_fast(fine, cfg.eval_every * fct)
        _, uf, vf, _ = filtered_fields(fine, fct)
        e = spectrum(uf, vf)
        track.append({"coarse_step": k, "spectrum": e, "energy": float(e[1:].sum()), "enstrophy": enstrophy(uf, vf)})
    log(f"closure study: reference track of {len(track)} filtered checkpoints")
    return start, track


def evaluate(cfg: StudyConfig, start, track, runners, device="cuda", log=print, spectra=None):
    """Coarse runs from the filtered start, scored against the reference track: mean |log E / E_ref| over the
    shells 1 .. n / 8 (well inside the filter), the relative energy and enstrophy errors.  `runners` maps a name
    to a function (steps) -> coarse Simulation after those steps (each keeps its own state).  `spectra` (a dict,
    optional) receives each runner's spectra and enstrophies at the checkpoints."""
    n = start[0].shape[0]
    kres = np.arange(1, n // 8 + 1)
    rows = []
    done = 0
    for ref in track:
        k = ref["coarse_step"]
        row = {"coarse_step": k, "enstrophy_ref": ref["enstrophy"]}
        for name, run in runners.items():
            sim = run(k - done)
            _, uc, vc = _coarse_uv(sim)
            e = spectrum(uc, vc)
            row[f"spectrum_err_{name}"] = float(np.mean(np.abs(np.log(e[kres] / ref["spectrum"][kres]))))
            row[f"energy_err_{name}"] = abs(float(e[1:].sum()) / ref["energy"] - 1.0)
            z = enstrophy(uc, vc)
            if spectra is not None:
                spectra.setdefault(name, []).append((e, z))
            row[f"enstrophy_{name}"] = z
            row[f"enstrophy_err_{name}"] = abs(z - ref["enstrophy"]) / ref["enstrophy"]
        done = k
        rows.append(row)
        log("  coarse step %d: " % k + ", ".join(f"{nm} spec {row[f'spectrum_err_{nm}']:.4f} ens "
                                                  f"{row[f'enstrophy_err_{nm}']:.4f}" for nm in runners))
    return rows


def time_averaged_errors(track, spectra, n):
    """Errors of the spectrum and the enstrophy averaged over the checkpoints after the start (the statistics of a
    chaotic flow, less sensitive to where its vortices happen to be than any one instant): mean over the shells
    1 .. n / 8 of |log <E> / <E_ref>|, its signed mean over the shells above n / 16 (negative: too little energy
    at small scales, i.e. too much dissipation), and |<Z> - <Z_ref>| / <Z_ref>."""
    kres = np.arange(1, n // 8 + 1)
    hi = kres >= n // 16
    R = np.mean([t["spectrum"] for t in track[1:]], 0)[kres]
    Z = float(np.mean([t["enstrophy"] for t in track[1:]]))
    out = {}
    for name, vals in spectra.items():
        M = np.mean([e for e, _ in vals[1:]], 0)[kres]
        out[f"spectrum_err_timeavg_{name}"] = float(np.mean(np.abs(np.log(M / R))))
        out[f"spectrum_log_bias_high_k_{name}"] = float(np.mean(np.log(M[hi] / R[hi])))
        out[f"enstrophy_err_timeavg_{name}"] = abs(float(np.mean([z for _, z in vals[1:]])) - Z) / Z
    return out


def summarize(rows, names):
    """Means over the checkpoints after the start, and the final checkpoint's errors, per runner."""
    out = {}
    for nm in names:
        out[f"spectrum_err_mean_{nm}"] = float(np.mean([r[f"spectrum_err_{nm}"] for r in rows[1:]]))
        out[f"enstrophy_err_mean_{nm}"] = float(np.mean([r[f"enstrophy_err_{nm}"] for r in rows[1:]]))
        out[f"energy_err_mean_{nm}"] = float(np.mean([r[f"energy_err_{nm}"] for r in rows[1:]]))
        out[f"spectrum_err_final_{nm}"] = rows[-1][f"spectrum_err_{nm}"]
        out[f"enstrophy_err_final_{nm}"] = rows[-1][f"enstrophy_err_{nm}"]
        out[f"energy_err_final_{nm}"] = rows[-1][f"energy_err_{nm}"]
    return out


def run_study(cfg: StudyConfig = StudyConfig(), device="cuda", log=print):
    """Build the training set from the fine run, train the MLP on the GPU, and compare the closure with
    plain Smagorinsky a posteriori over `eval_steps` coarse steps.  Returns (report, mlp, closure config)."""
    import time
    t_start = time.perf_counter()
    fine, data = build_dataset(cfg, device, log)
    mlp, ccfg, apriori, info = train_closure(cfg, data, device, log)
    del data
    start, track = reference_track(cfg, fine, log)
    del fine
    torch.cuda.empty_cache()
    rb, u, v = start
    smag = _coarse_sim(cfg, rb, u, v, cfg.cs_coarse, device)
    learned = ClosureSim(_coarse_sim(cfg, rb, u, v, cfg.cs_coarse, device), mlp, ccfg)

    def run_smag(k):
        _step_fast(smag, k)
        return smag

    def run_learned(k):
        learned.step(k)
        return learned.sim

    spectra = {}
    rows = evaluate(cfg, start, track, {"smagorinsky": run_smag, "learned": run_learned}, device, log, spectra)
    n = rb.shape[0]
    fb = int(learned.fallbacks.item())
    report = {
        "config": {k: getattr(cfg, k) for k in cfg.__dataclass_fields__},
        "coarse": n, **info, "apriori": apriori, "aposteriori": rows,
        "fallback_fraction": fb / float(max(1, cfg.eval_steps) * n * n),
        "seconds": round(time.perf_counter() - t_start, 1),
    }
    report["summary"] = {"spectrum_shells": f"1..{n // 8}", **summarize(rows, ("smagorinsky", "learned")),
                         **time_averaged_errors(track, spectra, n)}
    log(f"closure study: {report['summary']}")
    return report, mlp, ccfg
