"""GPU MLP kernels (BF16 matrix cores, FP32 accumulation) against the CPU FP64 MLP of oracle/mlp.py:
forward and gradient agree with the BF16-rounding mirror to FP32-accumulation error and with plain FP64 to
BF16 error; Adam is bit-identical to its FP32 restatement; the reduction is deterministic; training converges.

Tolerances: relative L2 error vs the mirror 2e-3 (FP32 summation order plus rare BF16 ties that round the other
way); vs plain FP64 3e-2 (about eight BF16 roundings of 2^-9 on the longest path)."""
import numpy as np
import pytest
import torch

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 18)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 3] * x[:, 7] - 0.3 * x[:, 11]).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    w /= w.sum()
    return x, t, w


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_pack_layout(gpu):
    from lbm.closure import MLP
    p = M.init(2)
    mlp = MLP(p)
    pk = mlp.packed.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    pk = pk.view(np.float32).astype(np.float64)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    w1t = pk[0:2048].reshape(64, 32)
    assert np.array_equal(w1t[:, :18], M.bf16(W1).T) and not w1t[:, 18:].any()
    assert np.array_equal(pk[2048:6144].reshape(64, 64), M.bf16(W2).T)
    assert np.array_equal(pk[6144:10240].reshape(64, 64), M.bf16(W2))
    w3t = pk[10240:].reshape(16, 64)
    assert np.array_equal(w3t[0], M.bf16(W3)) and not w3t[1:].any()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 257, 4096])
def test_forward_matches_cpu(gpu, n):
    from lbm.closure import MLP
    p = M.init(1)
    mlp = MLP(p)
    x, _, _ = _data(n, n)
    y = mlp.forward(_dev(x)).cpu().numpy()
    ym, _ = M.forward(p, x, mirror=True)
    y64, _ = M.forward(p, x)
    assert np.all(np.isfinite(y))
    assert _rel(y, ym) <= 2e-3, _rel(y, ym)
    assert _rel(y, y64) <= 3e-2, _rel(y, y64)


@pytest.mark.parametrize("n", [64, 1000, 8192])
def test_gradient_matches_cpu(gpu, n):
    from lbm.closure import MLP
    p = M.init(3)
    mlp = MLP(p)
    x, t, w = _data(n, 7 + n)
    g, loss = mlp.grad(_dev(x), _dev(t), _dev(w))
    g = g.cpu().numpy()
    gm, lm = M.grad(p, x, t, w, mirror=True)
    g64, l64 = M.grad(p, x, t, w)
    assert _rel(g, gm) <= 2e-3, _rel(g, gm)
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                   (M.OFF_W3, M.OFF_B3)):
        assert _rel(g[lo:hi], gm[lo:hi]) <= 5e-3, (lo, _rel(g[lo:hi], gm[lo:hi]))
    assert abs(g[M.OFF_B3] - gm[M.OFF_B3]) <= 2e-3 * (abs(gm[M.OFF_B3]) + 1e-3)
    assert _rel(g, g64) <= 3e-2, _rel(g, g64)
    assert abs(float(loss.item()) - lm) <= 2e-3 * lm
    g2, loss2 = mlp.grad(_dev(x), _dev(t), _dev(w))
    assert np.array_equal(g2.cpu().numpy(), g) and loss2.item() == loss.item()  # deterministic


def test_adam_step_bitwise(gpu):
    from lbm.closure import MLP
    p = M.init(4)
    mlp = MLP(p)
    x, t, w = _data(2048, 9)
    m0, v0 = np.zeros(M.NPARAM, np.float32), np.zeros(M.NPARAM, np.float32)
    for step in range(3):
        g = torch.empty_like(mlp.params)
        p_before = mlp.numpy()
        hp = mlp.train_step(_dev(x), _dev(t), _dev(w), lr=1e-3, grad=g)
        pc, m0, v0 = M.adam(p_before, m0, v0, g.cpu().numpy(), hp.lr, hp.beta1, hp.beta2, hp.eps, hp.bc1, hp.bc2)
        assert np.array_equal(mlp.numpy(), pc), f"step {step}"
        assert np.array_equal(mlp.m.cpu().numpy(), m0) and np.array_equal(mlp.v.cpu().numpy(), v0)
    # the repack inside the step matches a fresh pack of the new parameters
    fresh = MLP(mlp.numpy())
    assert torch.equal(fresh.packed, mlp.packed)


def test_training_converges(gpu):
    from lbm.closure import MLP
    mlp = MLP(M.init(5))
    x, t, w = _data(16384, 11)
    xd, td, wd = _dev(x), _dev(t), _dev(w)
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    mlp.train_step(xd, td, wd, lr=3e-3, loss=loss)
    first = loss.item()
    for _ in range(400):
        mlp.train_step(xd, td, wd, lr=3e-3, loss=loss)
    assert loss.item() < 0.1 * first, (first, loss.item())
