"""Writes the regression fixtures in this folder from the FP64 oracle (oracle/).

The reference holds no golden vectors, fixtures or tests for this path (it has no tests at all), so
these fixtures pin our own oracle against later changes; correctness itself is pinned by the closed-form
and analytic checks in tests/test_oracle.py.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "lbm-cuda_amd")]

import oracle as O  # noqa: E402
from lbm import geometry as G  # noqa: E402


def channel_32x16():
    W, H = 32, 16
    mask = G.channel_mask(W, H, cylinder=(9, 8, 3))
    prof = G.parabolic_profile(H, 0.05)
    f0 = G.channel_initial_state(W, H, prof, noise=1e-4, seed=1)
    tau_inv, cs, steps = float(np.float32(G.reynolds_tau_inv(0.05, 6, 20))), float(np.float32(0.1)), 10
    p = O.params(W, H, tau_inv, cs=cs)
    f = f0.copy()
    par = O.aa_steps(p, mask, f, prof, 0, steps)
    rho, ux, uy = O.macro(p, mask, f, prof, par)
    np.savez(os.path.join(HERE, "channel_32x16_mrt_les.npz"), mask=mask, inlet_u=prof, f0=f0, f_end=f, rho=rho,
             ux=ux, uy=uy, tau_inv=tau_inv, cs=cs, steps=steps)


if __name__ == "__main__":
    O.build()
    channel_32x16()
    print("fixtures written to", HERE)
