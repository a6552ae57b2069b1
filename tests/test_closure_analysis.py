"""CPU checks of the closure study's host-side analysis: box filter, subgrid-stress fit, spectrum, enstrophy
and stencil features, each against a closed form."""
import numpy as np

from lbm import closure as C


def test_box_filter_means_blocks():
    a = np.arange(64, dtype=np.float64).reshape(8, 8)
    b = C.box_filter(a, 4)
    assert b.shape == (2, 2)
    assert b[0, 0] == a[:4, :4].mean() and b[1, 1] == a[4:, 4:].mean()


def test_sgs_fit_recovers_eddy_viscosity():
    """tau_d = -2 nu S exactly gives nu* = nu everywhere the strain is non-zero."""
    n = 64
    x = np.arange(n)
    X, Y = np.meshgrid(x, x)
    u = 0.03 * np.sin(2 * np.pi * Y / n) + 0.01 * np.cos(2 * np.pi * 3 * X / n)
    v = 0.02 * np.cos(2 * np.pi * 2 * X / n)
    s1, s2 = C.strain(u, v)
    nu = 0.004
    tau = (-2 * nu * s1, 2 * nu * s1, -2 * nu * s2)  # txx - tyy = -4 nu s1, txy = -2 nu s2
    nu_star, w = C.sgs_target(u, v, tau)
    sel = w > 1e-12
    assert np.allclose(nu_star[sel], nu, rtol=1e-9)
    assert np.allclose(w, s1 * s1 + s2 * s2)


def test_spectrum_single_mode_and_parseval():
    n = 64
    x = np.arange(n)
    X, Y = np.meshgrid(x, x)
    u = 0.05 * np.cos(2 * np.pi * 5 * X / n)
    v = np.zeros_like(u)
    e = C.spectrum(u, v)
    assert np.argmax(e) == 5
    assert np.isclose(e.sum(), 0.5 * np.mean(u * u + v * v))  # Parseval (all shells up to n/2 hold it)


def test_enstrophy_of_shear_wave():
    n = 128
    y = np.arange(n)[:, None] * np.ones((1, n))
    k = 2 * np.pi * 3 / n
    u = 0.04 * np.sin(k * y)
    v = np.zeros_like(u)
    # omega = -du/dy = -0.04 k cos(k y) -> <omega^2> / 2 = (0.04 k)^2 / 4
    assert np.isclose(C.enstrophy(u, v), (0.04 * k) ** 2 / 4, rtol=1e-10)


def test_stencil_feature_order():
    n7 = np.arange(20, dtype=np.float64).reshape(4, 5)
    n8 = -n7
    f = C.stencil_features(np.stack([n7, n8]))
    assert f.shape == (20, 18)
    # cell (y=1, x=2): stencil s = 3 (dy+1) + (dx+1); feature 2 s + c
    row = f[1 * 5 + 2]
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            s = 3 * (dy + 1) + (dx + 1)
            assert row[2 * s] == n7[(1 + dy) % 4, (2 + dx) % 5]
            assert row[2 * s + 1] == n8[(1 + dy) % 4, (2 + dx) % 5]
    # periodic wrap at the corner cell (0, 0)
    assert f[0][0] == n7[3, 4]


def test_smagorinsky_nu_matches_rule():
    neq = np.array([[[3e-4]], [[-4e-4]]])
    tau0 = 0.52
    nu = C.smagorinsky_nu(neq, np.array([[1.0]]), tau0, 0.1)
    Q = np.sqrt(9e-8 + 16e-8 + 1e-8)
    tau_eff = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + 18 * 0.01 * Q))
    assert np.isclose(nu[0, 0], (tau_eff - tau0) / 3)


def test_windowed_fit_recovers_eddy_viscosity():
    n = 32
    x = np.arange(n)
    X, Y = np.meshgrid(x, x)
    u = 0.03 * np.sin(2 * np.pi * Y / n) + 0.01 * np.cos(2 * np.pi * 3 * X / n)
    v = 0.02 * np.cos(2 * np.pi * 2 * X / n)
    s1, s2 = C.strain(u, v)
    tau = (-2 * 0.003 * s1, 2 * 0.003 * s1, -2 * 0.003 * s2)
    nu_star, w = C.sgs_target(u, v, tau, window=5)
    assert np.allclose(nu_star, 0.003, rtol=1e-9)
    assert np.allclose(w, C.window_sum(s1 * s1 + s2 * s2, 5))
    assert np.isclose(C.window_sum(np.ones((n, n)), 5)[3, 7], 25.0)


def test_regularized_state_moments():
    """Mass and momentum are those of feq; the second moment carries Pi_neq = -(2/3) rho tau S."""
    from lbm import geometry as G
    n = 32
    x = np.arange(n)
    X, Y = np.meshgrid(x, x)
    rho = 1.0 + 0.01 * np.cos(2 * np.pi * X / n)
    u = 0.04 * np.sin(2 * np.pi * Y / n)
    v = 0.02 * np.cos(2 * np.pi * 2 * X / n)
    nu, cs = 1e-4, 0.1
    f = C.regularized_state(rho, u, v, nu, cs)
    feq = G.feq(rho, u, v)
    fn = f - feq
    assert np.allclose(fn.sum(0), 0, atol=1e-15)
    assert np.allclose((G.CX[:, None, None] * fn).sum(0), 0, atol=1e-15)
    s1, s2 = C.strain(u, v)
    dudx = 0.5 * (np.roll(u, -1, 1) - np.roll(u, 1, 1))
    dvdy = 0.5 * (np.roll(v, -1, 0) - np.roll(v, 1, 0))
    sxy = s2
    tau = 3 * (nu + cs * cs * np.sqrt(2 * (dudx ** 2 + dvdy ** 2 + 2 * sxy ** 2))) + 0.5
    pxy = (G.CX[:, None, None] * G.CY[:, None, None] * fn).sum(0)
    pxx_pyy = ((G.CX[:, None, None] ** 2 - G.CY[:, None, None] ** 2) * fn).sum(0)
    assert np.allclose(pxy, -(2 / 3) * rho * tau * sxy, rtol=1e-10, atol=1e-16)
    assert np.allclose(pxx_pyy, -(2 / 3) * rho * tau * (dudx - dvdy), rtol=1e-10, atol=1e-16)


def test_d4_augmentation_matches_transformed_fields():
    """The closure's symmetry augmentation is the stencil features of the rotated / reflected field.  A
    traceless symmetric tensor field (m7 ~ S_xx - S_yy, m8 ~ S_xy) rotated by 90 degrees, (x, y) -> (-y, x),
    becomes m7'(r) = -m7(R^T r), m8'(r) = -m8(R^T r); reflected, x -> -x, it becomes m7'(r) = m7(Mr),
    m8'(r) = -m8(Mr).  The features of the transformed field at r must be the augmented features of the original
    at the pre-image of r (periodic N x N grid, arrays indexed [y, x])."""
    from lbm import closure as C
    N = 12
    rng = np.random.default_rng(3)
    m7, m8 = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    X = C.stencil_features(np.stack([m7, m8]))  # [N*N, 18], row y * N + x
    T = np.zeros(N * N)
    Xa, _, _ = C.d4_augment(X.astype(np.float32), T, T)
    images = Xa.reshape(8, N * N, 18)
    yy, xx = np.mgrid[0:N, 0:N]
    # rotation: the value at r = (x, y) comes from R^T r = (y, -x)
    src_y, src_x = (-xx) % N, yy
    rot = C.stencil_features(np.stack([-m7[src_y, src_x], -m8[src_y, src_x]]))
    pre = (src_y * N + src_x).reshape(-1)
    # reflection: the value at (x, y) comes from (-x, y)
    ry, rx = yy, (-xx) % N
    refl = C.stencil_features(np.stack([m7[ry, rx], -m8[ry, rx]]))
    pre_r = (ry * N + rx).reshape(-1)
    # d4_augment's order: identity, reflection, rotation, rotation then reflection, ...
    assert np.allclose(images[0], X, atol=1e-6)
    assert np.allclose(images[2][pre], rot, atol=1e-6)
    assert np.allclose(images[1][pre_r], refl, atol=1e-6)
    # the 8 images are distinct group elements (no duplicate)
    assert len({img.tobytes() for img in images}) == 8
