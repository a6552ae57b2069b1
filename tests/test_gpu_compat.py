"""GPU: the reference-shaped kernel handles (lbm.compat) driven like lbm_smemaa_main.py / lbm_mine2.py /
lbm_smemaa_endfused_main.py, checked against the FP32 oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from lbm import compat
from lbm import geometry as G

pytestmark = pytest.mark.gpu

W, H = 256, 96


def _driver_state():
    """The reference drivers' initial state: rho = 1, two Gaussian ux blobs (lbm_smemaa_main.py:21-40)."""
    x = np.arange(W, dtype=np.float64)
    y = np.arange(H, dtype=np.float64)
    X, Y = np.meshgrid(x, y)
    t1 = np.exp(-((X - 60) ** 2 + (Y - 48) ** 2) / (2 * 10 ** 2))
    t2 = np.exp(-((X - 120) ** 2 + (Y - 48) ** 2) / (2 * 10 ** 2))
    ux = (0.1 * (t1 - t2)).astype(np.float32)
    f0 = G.feq(np.ones((H, W)), ux.astype(np.float64), 0.0).astype(np.float32)
    mask = G.channel_mask(W, H, cylinder=(180, 48, 8))
    return f0, mask


def _dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t if dtype is None else t.to(dtype)).cuda()


def test_smemaa_driver_flow():
    f0, mask = _driver_state()
    mod = compat.RawModule("lbm_core_smemaa.cu")
    k = mod.get_function("fused_lbmkernel")
    f_now, f_out = _dev(f0), torch.empty_like(_dev(f0))
    ux = torch.zeros((H, W), device="cuda")
    uy, rho = torch.zeros_like(ux), torch.zeros_like(ux)
    mask_d = _dev(mask, torch.bool)
    tau_inv, u_in = np.float32(1.7), np.float32(0.10)
    grid, block = ((W + 23) // 24, (H + 23) // 24), (32, 8)
    for _ in range(3):  # ping-pong as the captured graph does (lbm_smemaa_main.py:141-142)
        k(grid, block, (mask_d, f_now, f_out, ux, uy, rho, np.int32(W), np.int32(H), tau_inv, u_in))
        k(grid, block, (mask_d, f_out, f_now, ux, uy, rho, np.int32(W), np.int32(H), tau_inv, u_in))
    torch.cuda.synchronize()
    op = O.params(W, H, float(tau_inv), model=O.BGK)
    prof = G.cosine_ramp_profile(H, float(u_in)).astype(np.float32)
    ref = f0.copy()
    par = O.aa_steps(op, mask, ref, prof, 0, 24)
    assert par == 0
    sel = O.observable(op, mask, 0)
    got = f_now.cpu().numpy()
    assert np.array_equal(got[sel], ref[sel])
    # macros of the last launch are those of f_out's successor state == f_now here
    r, a, b = O.macro(op, mask, ref, prof, 0)
    upd = sel[0]
    assert np.array_equal(ux.cpu().numpy()[upd], a[upd]) and np.array_equal(rho.cpu().numpy()[upd], r[upd])


def test_core4_driver_flow():
    f0, mask = _driver_state()
    k = compat.RawModule("lbm_core4.cu").get_function("fused_lbmkernel")
    f_now, f_out = _dev(f0), _dev(f0)
    ux = torch.zeros((H, W), device="cuda")
    uy, rho = torch.zeros_like(ux), torch.zeros_like(ux)
    mask_d = _dev(mask, torch.bool)
    args = lambda a, b: (mask_d, a, b, ux, uy, rho, np.int32(W), np.int32(H), np.float32(1.8), np.float32(0.1),  # noqa
                         np.float32(1.8))
    for _ in range(5):
        k((64, 64), (32, 8), args(f_now, f_out))
        k((64, 64), (32, 8), args(f_out, f_now))
    torch.cuda.synchronize()
    op = O.params(W, H, 1.8, cs=0.2)
    prof = G.cosine_ramp_profile(H, float(np.float32(0.1))).astype(np.float32)
    a, b = f0.copy(), f0.copy()
    for _ in range(5):
        O.pull_step(op, mask, a, b, prof)
        O.pull_step(op, mask, b, a, prof)
    sel = O.observable(op, mask, 1)
    assert np.array_equal(f_now.cpu().numpy()[sel], a[sel])


def test_endfused_driver_flow_with_macro_and_visualize():
    f0, mask = _driver_state()
    mod = compat.RawModule("lbm_core_smemaa_policy.cu")
    k, mk, vk = (mod.get_function(n) for n in ("fused_lbmkernel", "macro_kernel", "visualizekernel"))
    f_now, f_out = _dev(f0), torch.empty_like(_dev(f0))
    mask_d = _dev(mask, torch.bool)
    for _ in range(2):
        k((1, 1), (32, 16), (mask_d, f_now, f_out, np.int32(W), np.int32(H), np.float32(1.0), np.float32(0.1)))
        k((1, 1), (32, 16), (mask_d, f_out, f_now, np.int32(W), np.int32(H), np.float32(1.0), np.float32(0.1)))
    ux = torch.zeros((H, W), device="cuda")
    uy, rho = torch.zeros_like(ux), torch.zeros_like(ux)
    mk((1, 1), (16, 16), (f_now, ux, uy, rho, mask_d, np.int32(W), np.int32(H), np.float32(0.1)))
    img = torch.zeros((H, W, 4), dtype=torch.uint8, device="cuda")
    vk((1, 1), (16, 16), (ux, uy, rho, img, mask_d, np.int32(W), np.int32(H), np.float32(50.0)))
    torch.cuda.synchronize()
    op = O.params(W, H, 1.0, model=O.BGK)
    prof = G.cosine_ramp_profile(H, float(np.float32(0.1))).astype(np.float32)
    ref = f0.copy()
    O.aa_steps(op, mask, ref, prof, 0, 24)
    sel = O.observable(op, mask, 0)
    assert np.array_equal(f_now.cpu().numpy()[sel], ref[sel])
    assert int(img[1:-1, 1:-1, 3].min()) == 255
