"""Independent numpy FP64 restatement of the reference kernels -- test infrastructure only.

Written separately from the C oracle, as a literal transcription of the reference's formulas in the
reference's own order, so that a transcription slip in either one shows up as a disagreement:

- ``pull_step``: ``fused_lbmkernel`` of lbm_core4.cu:7-212 (pull + half-way bounce-back, moments with
  clamps, MRT + Smagorinsky, Zou-He inlet with inlet MRT, population clamp, W-2 -> W-1 outlet copy), with
  the inlet speed given per row and Cs/s-rates as arguments instead of the literals at :6,:123.
- ``pull_step_bgk``: the same streaming with the BGK collision of lbm_core_smemaa.cu:194-220.

Vectorised over cells; wall rows, solids and the outlet ghost keep their previous values, as in the
reference (f_out is only written for updated cells; the caller passes f_out = f_now.copy()).
"""
from __future__ import annotations

import numpy as np

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
WQ = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])


def _gather(f_now, mask, periodic):
    """f_q(x) = mask[x - c_q] ? f_now[opp q](x) : f_now[q](x - c_q)   (lbm_core4.cu:129-134).
    Out-of-domain sources (only x = -1 at the inlet) return 0 (lbm_core4.cu:36-37)."""
    _, H, W = f_now.shape
    g = np.zeros_like(f_now)
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    for q in range(9):
        sx, sy = xx - CX[q], yy - CY[q]
        if periodic:
            sx, sy = sx % W, sy % H
            inside = np.ones_like(sx, dtype=bool)
        else:
            inside = (sx >= 0) & (sx < W) & (sy >= 0) & (sy < H)
        sxc, syc = np.clip(sx, 0, W - 1), np.clip(sy, 0, H - 1)
        solid = mask[syc, sxc].astype(bool)
        val = np.where(solid, f_now[OPP[q]], f_now[q][syc, sxc])
        g[q] = np.where(inside, val, 0.0)
    return g


def _roles(mask, periodic):
    H, W = mask.shape
    fluid = ~mask.astype(bool)
    inlet = np.zeros((H, W), bool)
    if not periodic:
        fluid[0, :] = False
        fluid[H - 1, :] = False
        fluid[:, W - 1] = False
        fluid[:, 0] = False
        inlet[1:H - 1, 0] = True
    return fluid, inlet


def _zou_he(f, u):
    """lbm_core4.cu:54-57 on arrays (f is [9, n], u is [n])."""
    f = f.copy()
    rho = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - u)
    f[1] = f[3] + (2.0 / 3.0) * rho * u
    f[5] = f[7] - 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * rho * u
    f[8] = f[6] + 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * rho * u
    return f, rho


def _mrt(f, rho, ux, uy, omega_stress, s_e, s_eps, s_q):
    """lbm_core4.cu:150-199 with the relaxation rate of m7/m8 given per cell."""
    f14 = f[1] + f[2] + f[3] + f[4]
    f58 = f[5] + f[6] + f[7] + f[8]
    m = [None] * 9
    m[0] = rho
    m[1] = -4.0 * f[0] + 2.0 * f58 - f14
    m[2] = 4.0 * f[0] + f58 - 2.0 * f14
    m[3] = rho * ux
    m[4] = -2.0 * (f[1] - f[3]) + f[5] - f[6] - f[7] + f[8]
    m[5] = rho * uy
    m[6] = -2.0 * (f[2] - f[4]) + f[5] + f[6] - f[7] - f[8]
    m[7] = f[1] - f[2] + f[3] - f[4]
    m[8] = f[5] - f[6] + f[7] - f[8]
    usq = ux * ux + uy * uy
    m1eq = (3.0 * usq - 2.0) * rho
    m2eq = (1.0 - 3.0 * usq) * rho
    m7eq = rho * (ux * ux - uy * uy)
    m8eq = rho * ux * uy
    m[1] = m[1] - s_e * (m[1] - m1eq)
    m[2] = m[2] - s_eps * (m[2] - m2eq)
    m[4] = m[4] - s_q * (m[4] + m[3])
    m[6] = m[6] - s_q * (m[6] + m[5])
    m[7] = m[7] - omega_stress * (m[7] - m7eq)
    m[8] = m[8] - omega_stress * (m[8] - m8eq)
    t0 = m[0] - m[1] + m[2]
    t1 = 4.0 * m[0] - m[1] - 2.0 * m[2]
    t2 = 4.0 * m[0] + 2.0 * m[1] + m[2]
    out = np.empty_like(f)
    out[0] = t0 / 9.0
    out[1] = (t1 + 6.0 * m[3] - 6.0 * m[4] + 9.0 * m[7]) / 36.0
    out[2] = (t1 + 6.0 * m[5] - 6.0 * m[6] - 9.0 * m[7]) / 36.0
    out[3] = (t1 - 6.0 * m[3] + 6.0 * m[4] + 9.0 * m[7]) / 36.0
    out[4] = (t1 - 6.0 * m[5] + 6.0 * m[6] - 9.0 * m[7]) / 36.0
    out[5] = (t2 + 6.0 * m[3] + 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] + 9.0 * m[8]) / 36.0
    out[6] = (t2 - 6.0 * m[3] - 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] - 9.0 * m[8]) / 36.0
    out[7] = (t2 - 6.0 * m[3] - 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] + 9.0 * m[8]) / 36.0
    out[8] = (t2 + 6.0 * m[3] + 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] - 9.0 * m[8]) / 36.0
    return np.minimum(10.0, np.maximum(0.0, out))


def _moments(g, clamp):
    """rho, ux, uy of gathered populations g [9, n] with the clamps of lbm_core4.cu:139-144."""
    rho = g.sum(axis=0)
    ux = (g * CX[:, None]).sum(axis=0) / rho
    uy = (g * CY[:, None]).sum(axis=0) / rho
    if clamp:
        ux = np.minimum(0.4, np.maximum(-0.4, ux))
        uy = np.minimum(0.4, np.maximum(-0.4, uy))
        rho = np.minimum(1.5, np.maximum(0.0, rho))
    return rho, ux, uy


def _neq(g, rho, ux, uy):
    """Non-equilibrium parts of the stress moments (lbm_core4.cu:172-173): m7 - m7eq, m8 - m8eq."""
    m7 = g[1] - g[2] + g[3] - g[4]
    m8 = g[5] - g[6] + g[7] - g[8]
    return m7 - rho * (ux * ux - uy * uy), m8 - rho * ux * uy


def neq(f, mask, parity, clamp=True, periodic=True):
    """m7neq, m8neq [2, H, W] of every fluid cell's pre-collision populations (0 elsewhere), for an AA state
    ``f`` of the given parity: canonical (0) holds the pre-collision populations in place; reversed (1) holds
    post-collision ones in opposite slots, f[opp q] = f*_q, which stream in as a pull step gathers them."""
    _, H, W = f.shape
    g = f if parity == 0 else _gather(f[OPP], mask, periodic)
    fluid, _ = _roles(mask, periodic)
    out = np.zeros((2, H, W))
    gi = g[:, fluid]
    rho, ux, uy = _moments(gi, clamp)
    out[0][fluid], out[1][fluid] = _neq(gi, rho, ux, uy)
    return out


def pull_step(f_now, mask, inlet_u, tau_inv, cs, s_e=1.1, s_eps=1.1, s_q=1.1, clamp=True, periodic=False):
    """One lbm_core4 step; returns (f_out, rho, ux, uy).  Macro arrays are 0 where not written."""
    _, H, W = f_now.shape
    g = _gather(f_now, mask, periodic)
    fluid, inlet = _roles(mask, periodic)
    f_out = f_now.copy()
    rho_o, ux_o, uy_o = (np.zeros((H, W)) for _ in range(3))

    # interior fluid, lbm_core4.cu:124-199
    gi = g[:, fluid]
    rho, ux, uy = _moments(gi, clamp)
    rho_o[fluid], ux_o[fluid], uy_o[fluid] = rho, ux, uy
    omega = np.full(rho.shape, float(tau_inv))
    if cs != 0.0:
        m7n, m8n = _neq(gi, rho, ux, uy)
        Q = np.sqrt(m7n * m7n + m8n * m8n + 1e-8)
        tau0 = 1.0 / tau_inv
        tau_eff = 0.5 * (tau0 + np.sqrt(tau0 * tau0 + 18.0 * cs * cs * Q / rho))
        omega = 1.0 / tau_eff
    f_out[:, fluid] = _mrt(gi, rho, ux, uy, omega, s_e, s_eps, s_q)

    # inlet, lbm_core4.cu:29-118 (Zou-He, MRT without LES or clamps, rate tau_inv)
    if inlet.any():
        ys = np.nonzero(inlet[:, 0])[0]
        u = inlet_u[ys]
        gz, rz = _zou_he(g[:, ys, 0], u)
        rho_o[ys, 0], ux_o[ys, 0], uy_o[ys, 0] = rz, u, 0.0
        f_out[:, ys, 0] = _mrt(gz, rz, u, np.zeros_like(u), float(tau_inv), s_e, s_eps, s_q)

    if not periodic:  # outlet ghost copy, lbm_core4.cu:114-116,207-209
        upd = fluid[:, W - 2] | inlet[:, W - 2]
        f_out[:, upd, W - 1] = f_out[:, upd, W - 2]
    return f_out, rho_o, ux_o, uy_o


def pull_step_bgk(f_now, mask, inlet_u, tau_inv, periodic=False):
    """Pull streaming with the BGK collision of lbm_core_smemaa.cu:194-220 (no u/rho clamps)."""
    _, H, W = f_now.shape
    g = _gather(f_now, mask, periodic)
    fluid, inlet = _roles(mask, periodic)
    f_out = f_now.copy()

    def bgk(f, rho, ux, uy):
        usq = ux * ux + uy * uy
        out = np.empty_like(f)
        for q in range(9):
            cu = CX[q] * ux + CY[q] * uy
            feq = WQ[q] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
            out[q] = f[q] + tau_inv * (feq - f[q])
        return np.minimum(10.0, np.maximum(0.0, out))

    gi = g[:, fluid]
    rho = gi.sum(axis=0)
    ux = (gi * CX[:, None]).sum(axis=0) / rho
    uy = (gi * CY[:, None]).sum(axis=0) / rho
    f_out[:, fluid] = bgk(gi, rho, ux, uy)
    if inlet.any():
        ys = np.nonzero(inlet[:, 0])[0]
        u = inlet_u[ys]
        gz, rz = _zou_he(g[:, ys, 0], u)
        f_out[:, ys, 0] = bgk(gz, rz, u, np.zeros_like(u))
    if not periodic:
        upd = fluid[:, W - 2] | inlet[:, W - 2]
        f_out[:, upd, W - 1] = f_out[:, upd, W - 2]
    return f_out
