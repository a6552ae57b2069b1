"""Masks, inlet profiles and initial states, built in FP64 on the host.

The reference builds its mask from an image (``lbm_smemaa_main.py:53-68``) and then forces the boundary
roles: rows 0 and H-1 solid (``:61-62``), the inlet column x=0 and the outlet ghost column x=W-1 open on
rows 1..H-2 (``:65-66``).  ``channel_mask`` applies the same forcing to an analytic geometry;
``finalize_channel_mask`` applies it to any user mask (for example one read from a BMP by
``lbm.io.mask_from_bmp``).
"""
from __future__ import annotations

import math

import numpy as np

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.int64)
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.int64)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int64)
WEIGHTS = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float64)


def finalize_channel_mask(mask):
    """Force the channel roles of lbm_smemaa_main.py:61-66 and check the outlet column.

    Returns a C-contiguous uint8 [H, W] copy.  Raises if a solid cell sits on column W-2 (the cell that
    feeds the zero-gradient ghost must be fluid)."""
    m = np.array(mask, dtype=np.uint8, copy=True)
    if m.ndim != 2:
        raise ValueError("mask must be 2-D [H, W]")
    H, W = m.shape
    if H < 3 or W < 3:
        raise ValueError(f"channel lattice needs W, H >= 3, got {W}x{H}")
    m[m != 0] = 1
    m[0, :] = 1
    m[H - 1, :] = 1
    m[1:H - 1, 0] = 0
    m[1:H - 1, W - 1] = 0
    if m[1:H - 1, W - 2].any():
        raise ValueError("solid cells on column W-2 are not supported (it feeds the outlet ghost column)")
    return np.ascontiguousarray(m)


def channel_mask(W, H, cylinder=None):
    """Channel with wall rows and an optional solid disc ``cylinder=(cx, cy, r)`` (cells with
    (x-cx)^2 + (y-cy)^2 <= r^2 at integer cell coordinates)."""
    m = np.zeros((H, W), np.uint8)
    if cylinder is not None:
        cx, cy, r = cylinder
        yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
        m[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = 1
    return finalize_channel_mask(m)


def wake_window(W, H, cx, cy, R, tile=(128, 12)):
    """Base window (x0, y0, w, h) over the near wake of a body of half-height R centred at (cx, cy) in a channel
    flowing in +x: the region [cx, cx + 4 R) x [cy - 2 R, cy + 2 R), where FP16 storage relative to the inlet
    profile keeps most of its error.  The window is the inside of the whole tiles of the two-step kernel's tiling
    (`tile` = columns x rows of a tile) that this region holds, less one cell at each side, so that only those tiles
    take the kernel's per-cell form (a tile whose halo reaches a window cell takes it too); a region holding no
    whole tile (small lattices) is taken as it is.  Clipped to 1 <= x <= W - 3, 1 <= y <= H - 2."""
    tw, th = tile
    rx0, rx1, ry0, ry1 = cx, cx + 4 * R, cy - 2 * R, cy + 2 * R
    tx0, tx1 = -(-rx0 // tw), rx1 // tw
    ty0, ty1 = -(-ry0 // th), ry1 // th
    if tx1 > tx0 and ty1 > ty0:
        x0, x1, y0, y1 = tx0 * tw + 1, tx1 * tw - 1, ty0 * th + 1, ty1 * th - 1
    else:
        x0, x1, y0, y1 = rx0, rx1, ry0, ry1
    x0, y0 = max(1, x0), max(1, y0)
    x1, y1 = min(W - 2, x1), min(H - 1, y1)
    if x1 <= x0 or y1 <= y0:
        return None
    return (x0, y0, x1 - x0, y1 - y0)


def obstacle_wake_window(mask, tile=(128, 12)):
    """wake_window of the bodies in a channel mask [H, W] (solid cells off the wall rows): centred on their
    bounding box, R its half-height; None without a body."""
    m = np.asarray(mask)
    H, W = m.shape
    ys, xs = np.nonzero(m[1:-1])
    if ys.size == 0:
        return None
    ys = ys + 1
    cx, cy = (int(xs.min()) + int(xs.max())) // 2, (int(ys.min()) + int(ys.max())) // 2
    R = max(1, (int(ys.max()) - int(ys.min())) // 2)
    return wake_window(W, H, cx, cy, R, tile)


def periodic_mask(W, H):
    return np.zeros((H, W), np.uint8)


def parabolic_profile(H, u_max):
    """Poiseuille profile with half-way walls at y = 1/2 and y = H - 3/2:
    u(y) = 4 u_max (y - 1/2)(H - 3/2 - y) / (H - 2)^2 for rows 1..H-2, 0 on the wall rows."""
    y = np.arange(H, dtype=np.float64)
    u = 4.0 * u_max * (y - 0.5) * (H - 1.5 - y) / float(H - 2) ** 2
    u[0] = 0.0
    u[H - 1] = 0.0
    return u


def cosine_ramp_profile(H, u_in, width=50.0):
    """The reference's inlet speed u_in * s(y), s = 1/2 (1 - cos(3.14159 d / width)) for d < width, with
    d = min(y, H-1-y)   (lbm_core4.cu:46-51), evaluated in FP64."""
    y = np.arange(H, dtype=np.float64)
    d = np.minimum(y, H - 1 - y)
    s = np.where(d < width, 0.5 * (1.0 - np.cos(3.14159 * d / width)), 1.0)
    return u_in * s


def feq(rho, ux, uy):
    """w_q rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) in FP64, shape [9, ...]   (lbm_smemaa_main.py:35-40)."""
    rho = np.asarray(rho, np.float64)
    ux = np.asarray(ux, np.float64)
    uy = np.asarray(uy, np.float64)
    shape = np.broadcast(rho, ux, uy).shape
    usq = ux * ux + uy * uy
    out = np.empty((9,) + shape)
    for q in range(9):
        cu = CX[q] * ux + CY[q] * uy
        out[q] = WEIGHTS[q] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    return out


def channel_initial_state(W, H, profile, noise=1e-4, seed=0):
    """rho = 1, ux = inlet profile on every column, uy ~ U(-noise, noise) from default_rng(seed); f = feq."""
    rng = np.random.default_rng(seed)
    ux = np.broadcast_to(np.asarray(profile, np.float64)[:, None], (H, W))
    uy = rng.uniform(-noise, noise, size=(H, W)) if noise else np.zeros((H, W))
    return feq(np.ones((H, W)), ux, uy)


def reynolds_tau_inv(u, length, re):
    """tau_inv = 1/tau with nu = u L / Re and tau = 3 nu + 1/2 (lattice units)."""
    nu = u * length / re
    return 1.0 / (3.0 * nu + 0.5)


def nu_tau_inv(nu):
    return 1.0 / (3.0 * nu + 0.5)


def shear_layer_velocity(W, H, u0=0.04, delta=80.0, amp=0.002, noise=1e-4, seed=3):
    """Periodic double shear layer on cell centres (j + 1/2):
    ux = u0 tanh(delta (y/H - 1/4)) for y < H/2, u0 tanh(delta (3/4 - y/H)) otherwise;
    uy = amp sin(2 pi x / W); plus U(-noise, noise) on both components."""
    rng = np.random.default_rng(seed)
    yc = (np.arange(H, dtype=np.float64) + 0.5) / H
    xc = np.arange(W, dtype=np.float64)
    ux1 = np.where(yc < 0.5, u0 * np.tanh(delta * (yc - 0.25)), u0 * np.tanh(delta * (0.75 - yc)))
    ux = np.broadcast_to(ux1[:, None], (H, W)).copy()
    uy = np.broadcast_to(amp * np.sin(2.0 * math.pi * xc / W)[None, :], (H, W)).copy()
    if noise:
        ux += rng.uniform(-noise, noise, size=(H, W))
        uy += rng.uniform(-noise, noise, size=(H, W))
    return ux, uy


def taylor_green_velocity(W, H, u0, t=0.0, nu=0.0, mx=1, my=1):
    """Decaying Taylor-Green vortex array with kx = 2 pi mx / W, ky = 2 pi my / H (mx, my whole periods across
    the lattice): ux = -u0 (ky/k) cos(kx x) sin(ky y) e^(-nu k^2 t), uy = u0 (kx/k) sin(kx x) cos(ky y) e^(-nu k^2 t),
    k^2 = kx^2 + ky^2 -- divergence free, amplitude u0 at t = 0."""
    kx, ky = 2.0 * math.pi * mx / W, 2.0 * math.pi * my / H
    x = np.arange(W, dtype=np.float64)[None, :]
    y = np.arange(H, dtype=np.float64)[:, None]
    k = math.hypot(kx, ky)
    decay = math.exp(-nu * (kx * kx + ky * ky) * t)
    ux = -u0 * (ky / k) * np.cos(kx * x) * np.sin(ky * y) * decay
    uy = u0 * (kx / k) * np.sin(kx * x) * np.cos(ky * y) * decay
    return np.broadcast_to(ux, (H, W)).copy(), np.broadcast_to(uy, (H, W)).copy()


def taylor_green_density(W, H, u0, mx=1, my=1):
    """rho = 1 + 3 p of the Taylor-Green pressure p = -(A^2 cos(2 kx x) + B^2 cos(2 ky y)) / 4, A = u0 ky / k,
    B = u0 kx / k (the velocity of taylor_green_velocity with the same modes); starting from it avoids the
    acoustic transient of a uniform density."""
    kx, ky = 2.0 * math.pi * mx / W, 2.0 * math.pi * my / H
    k = math.hypot(kx, ky)
    a, b = u0 * ky / k, u0 * kx / k
    x = np.arange(W, dtype=np.float64)[None, :]
    y = np.arange(H, dtype=np.float64)[:, None]
    return 1.0 - 0.75 * (a * a * np.cos(2 * kx * x) + b * b * np.cos(2 * ky * y))


def random_phase_velocity(W, H, k0, u_rms, seed, power=4.0):
    """Solenoidal random-phase field from a stream function with E(k) ~ k^power exp(-2 (k/k0)^2),
    rescaled to the given u_rms (mean of ux^2 + uy^2 over the box, square-rooted)."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(W) * W
    ky = np.fft.fftfreq(H) * H
    KX, KY = np.meshgrid(kx, ky)
    k = np.hypot(KX, KY)
    with np.errstate(divide="ignore", invalid="ignore"):
        # |psi_k|^2 ~ E(k) / (k^3) in 2-D (energy shell 2 pi k |u_k|^2, u = k psi)
        amp = np.where(k > 0, np.sqrt(k ** power * np.exp(-2.0 * (k / k0) ** 2) / np.maximum(k, 1e-30) ** 3), 0.0)
    phase = np.exp(2j * math.pi * rng.random((H, W)))
    psi_k = amp * phase
    psi = np.real(np.fft.ifft2(psi_k))
    psi_k = np.fft.fft2(psi)
    ux = np.real(np.fft.ifft2(1j * 2 * math.pi * KY / H * psi_k))
    uy = np.real(np.fft.ifft2(-1j * 2 * math.pi * KX / W * psi_k))
    scale = u_rms / math.sqrt(float(np.mean(ux * ux + uy * uy)))
    return ux * scale, uy * scale


# ------------------------------------------------------------------------------------------------ image I/O
def read_bmp(path):
    """Pixels of an uncompressed 24- or 32-bit BMP as uint8 [rows, cols, 3] in B, G, R order, top row first
    (the channel order cv2.imread returns)."""
    import struct
    data = open(path, "rb").read()
    if data[:2] != b"BM":
        raise ValueError(f"{path}: not a BMP file")
    off = struct.unpack_from("<I", data, 10)[0]
    width, height = struct.unpack_from("<ii", data, 18)
    bpp, comp = struct.unpack_from("<HI", data, 28)
    if bpp not in (24, 32) or comp not in (0, 3):
        raise ValueError(f"{path}: only uncompressed 24/32-bit BMP is supported (bpp {bpp}, compression {comp})")
    nb = bpp // 8
    stride = (width * nb + 3) & ~3
    rows = abs(height)
    px = np.frombuffer(data, np.uint8, count=stride * rows, offset=off).reshape(rows, stride)
    img = px[:, :width * nb].reshape(rows, width, nb)[:, :, :3]
    return np.ascontiguousarray(img[::-1] if height > 0 else img)  # positive height: stored bottom-up


def resize_linear(img, W, H):
    """Bilinear resize to H x W with half-pixel centres (src = (dst + 1/2) scale - 1/2, edges clamped), rounded
    to uint8 -- the geometry of cv2.resize's default INTER_LINEAR (its fixed-point weights are not reproduced)."""
    h, w = img.shape[:2]
    if (h, w) == (H, W):
        return img.copy()

    def axis(n_out, n_in):
        s = (np.arange(n_out) + 0.5) * (n_in / n_out) - 0.5
        s = np.clip(s, 0, n_in - 1)
        i0 = np.floor(s).astype(int)
        i1 = np.minimum(i0 + 1, n_in - 1)
        return i0, i1, s - i0

    y0, y1, fy = axis(H, h)
    x0, x1, fx = axis(W, w)
    a = img.astype(np.float64)
    top = a[y0][:, x0] * (1 - fx)[None, :, None] + a[y0][:, x1] * fx[None, :, None]
    bot = a[y1][:, x0] * (1 - fx)[None, :, None] + a[y1][:, x1] * fx[None, :, None]
    out = top * (1 - fy)[:, None, None] + bot * fy[:, None, None]
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def mask_from_image(path, W, H):
    """Channel mask from an image, as the reference's load_mask_from_image (lbm_smemaa_main.py:53-68): resize
    to W x H, solid where the green channel is 0, rows 0 and H-1 solid, columns 0 and W-1 open on rows 1..H-2."""
    img = resize_linear(read_bmp(path), W, H)
    mask = (img[:, :, 1] == 0).astype(np.uint8)
    return finalize_channel_mask(mask)


def write_ppm(path, rgba):
    """Binary PPM (P6) of an [H, W, 4] or [H, W, 3] uint8 image (the lbm_visualize output), row 0 first."""
    img = np.ascontiguousarray(np.asarray(rgba, np.uint8)[:, :, :3])
    H, W = img.shape[:2]
    with open(path, "wb") as fh:
        fh.write(f"P6\n{W} {H}\n255\n".encode())
        fh.write(img.tobytes())
