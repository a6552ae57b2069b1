"""The workloads this engine is measured and checked on, as builders of (parameters, mask, inlet, f0).

Channel flow past a cylinder (the reference's only scenario, lbm_smemaa_main.py:53-71 with an image mask;
here the disc is analytic):
  cylinder_small  512 x 128, R = 16 at (128, 64), Re = 100, Cs = 0.10, 1000 steps, seed 1
  cylinder_large  4096 x 1024, R = 64 at (1024, 512), Re = 1000, Cs = 0.15, 10000 steps, seed 2
Both: wall rows, parabolic velocity inlet with u_max = 0.05, zero-gradient outlet, MRT + Smagorinsky,
rho = 1 and ux = inlet profile everywhere at t = 0, uy ~ U(-1e-4, 1e-4); Re = u_max D / nu with D = 2R.
Periodic double shear layer 8192 x 2048, nu = 1e-4, MRT without LES, seed 3 (precision study).
Ensemble of 64 periodic lattices, W and H each drawn uniformly from {256, 384, 512, 768, 1024, 1536, 2048},
nu log-uniform in [1e-4, 1e-2], random-phase solenoidal field E(k) ~ k^4 exp(-2 (k/8)^2) at u_rms = 0.04,
MRT + Smagorinsky Cs = 0.1, rho = 1, seed 4 (batched-ensemble study).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import geometry as geo


@dataclass
class Case:
    name: str
    W: int
    H: int
    tau_inv: float
    cs: float
    model: str
    bc: str
    mask: np.ndarray
    inlet_u: np.ndarray
    f0: np.ndarray
    steps: int
    meta: dict = field(default_factory=dict)

    def sim_kwargs(self):
        return dict(cs=self.cs, model=self.model, bc=self.bc, mask=self.mask, inlet_u=self.inlet_u, f0=self.f0)


def cylinder(W, H, R, cx, cy, re, cs, steps, seed, u_max=0.05, name="cylinder"):
    D = 2 * R
    tau_inv = float(np.float32(geo.reynolds_tau_inv(u_max, D, re)))
    mask = geo.channel_mask(W, H, cylinder=(cx, cy, R))
    prof = geo.parabolic_profile(H, u_max)
    f0 = geo.channel_initial_state(W, H, prof, noise=1e-4, seed=seed)
    return Case(name, W, H, tau_inv, float(np.float32(cs)), "mrt", "channel", mask, prof, f0, steps,
                dict(u_ref=u_max, D=D, re=re, nu=u_max * D / re, cx=cx, cy=cy, R=R))


def wake_window(case, tile=(128, 12)):
    """Base window (x0, y0, w, h) over a cylinder case's near wake (geometry.wake_window of its cylinder): at
    4096 x 1024 and 10 000 steps, 9/10 of the FP16 row-based storage's squared velocity error lies in
    [960, 1472) x [384, 640), over half of it in the first 128 columns behind the body."""
    cx, cy, R = (int(case.meta[k]) for k in ("cx", "cy", "R"))
    return geo.wake_window(case.W, case.H, cx, cy, R, tile)


def cylinder_small():
    return cylinder(512, 128, 16, 128, 64, 100.0, 0.10, 1000, seed=1, name="cylinder_512x128")


def cylinder_large():
    return cylinder(4096, 1024, 64, 1024, 512, 1000.0, 0.15, 10000, seed=2, name="cylinder_4096x1024")


def cylinder_scaled(W, H, steps=100, seed=2):
    """The large cylinder case resized (same Re, Cs, relative geometry) -- for quick tests."""
    R = max(2, H // 16)
    return cylinder(W, H, R, W // 4, H // 2, 1000.0, 0.15, steps, seed=seed, name=f"cylinder_{W}x{H}")


def shear_layer(W=8192, H=2048, nu=1e-4, steps=20000, seed=3):
    ux, uy = geo.shear_layer_velocity(W, H, seed=seed)
    f0 = geo.feq(np.ones((H, W)), ux, uy)
    tau_inv = float(np.float32(geo.nu_tau_inv(nu)))
    return Case(f"shear_layer_{W}x{H}", W, H, tau_inv, 0.0, "mrt", "periodic", geo.periodic_mask(W, H),
                np.zeros(H), f0, steps, dict(nu=nu))


ENSEMBLE_SIZES = (256, 384, 512, 768, 1024, 1536, 2048)


@dataclass
class Member:
    """One lattice of an ensemble: size, rate, and a builder of its initial populations (built on demand,
    so that a 64-member ensemble never holds every FP64 state on the host at once)."""
    W: int
    H: int
    nu: float
    tau_inv: float
    cs: float
    seed: int
    k0: float = 8.0
    u_rms: float = 0.04

    def f0(self, dtype=np.float64):
        """Equilibrium populations at rho = 1 and the member's velocity field, computed in FP64."""
        ux, uy = geo.random_phase_velocity(self.W, self.H, k0=self.k0, u_rms=self.u_rms, seed=self.seed)
        return geo.feq(np.ones((self.H, self.W)), ux, uy).astype(dtype, copy=False)


def ensemble(n=64, sizes=ENSEMBLE_SIZES, nu_range=(1e-4, 1e-2), cs=0.1, seed=4):
    """Members of the batched-ensemble workload (periodic, MRT + Smagorinsky)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W = int(rng.choice(sizes))
        H = int(rng.choice(sizes))
        nu = float(np.exp(rng.uniform(np.log(nu_range[0]), np.log(nu_range[1]))))
        out.append(Member(W, H, nu, float(np.float32(geo.nu_tau_inv(nu))), float(np.float32(cs)),
                          seed=seed * 1000 + i))
    return out
