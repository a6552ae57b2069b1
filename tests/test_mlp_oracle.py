"""CPU pins of the FP64 MLP restatement used to check the GPU kernels: the analytic gradient against central
finite differences, GELU' against differences of GELU, the BF16 rounding, and the parameter layout."""
import numpy as np
import pytest

from oracle import mlp as M


def test_parameter_count():
    assert M.NPARAM == 18 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 5441


def test_gelu_derivative():
    z = np.linspace(-6, 6, 121)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_d(z))) < 1e-8
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([10.0]))[0] - 10.0) < 1e-12  # gelu(z) -> z for large z


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert list(M.bf16(x)) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0]


@pytest.mark.parametrize("mirror", [False, True])
def test_gradient_matches_finite_differences(mirror):
    rng = np.random.default_rng(3)
    p = M.init(5).astype(np.float64) * 0.5
    x = rng.normal(0, 1, (23, 18))
    t = rng.normal(0, 1, 23)
    w = rng.uniform(0.2, 1.0, 23)
    g, loss = M.grad(p, x, t, w)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 40), np.arange(M.OFF_B1, M.OFF_B1 + 8),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 40), np.arange(M.OFF_B2, M.OFF_B2 + 8),
                          np.arange(M.OFF_W3, M.OFF_B3), [M.OFF_B3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.grad(pp, x, t, w)[1] - M.grad(pm, x, t, w)[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(g[i])), (i, fd, g[i])
    if mirror:
        # the BF16 mirror stays close to the exact gradient (a few BF16 roundings deep)
        gm, lm = M.grad(p, x, t, w, mirror=True)
        assert np.linalg.norm(gm - g) <= 3e-2 * np.linalg.norm(g)
        assert abs(lm - loss) <= 3e-2 * loss


def test_adam_first_step_moves_by_lr():
    p = np.zeros(4, np.float32)
    g = np.array([1.0, -2.0, 1e-3, 0.0], np.float32)
    p1, m1, v1 = M.adam(p, np.zeros(4), np.zeros(4), g, 1e-2, 0.9, 0.999, 1e-8, 1 - 0.9, 1 - 0.999)
    # step 1 with bias correction: update = lr * g / (|g| + eps) ~= lr * sign(g)
    assert np.allclose(p1[:3], -1e-2 * np.sign(g[:3]), rtol=1e-4)
    assert p1[3] == 0.0
