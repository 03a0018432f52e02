"""CPU checks that pin the closure's FP64 restatement (oracle/closure.py)."""
import numpy as np
import pytest

from oracle import closure as C
from oracle import oracle as O
from tests.helpers import random_state


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def _net(seed=0):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(C.NW) * 0.3
    norm = (rng.standard_normal(C.NIN) * 0.1, 1.0 + rng.random(C.NIN), 0.5 + rng.random(C.NOUT))
    X = rng.standard_normal((40, C.NIN))
    Y = rng.standard_normal((40, C.NOUT))
    return w, norm, X, Y


def test_gradient_matches_finite_differences():
    w, norm, X, Y = _net()
    loss, g = C.loss_and_grad(w, norm, X, Y)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.integers(C.OFF_W1, C.OFF_B1, 4), rng.integers(C.OFF_B1, C.OFF_W2, 2),
                          rng.integers(C.OFF_W2, C.OFF_B2, 6), rng.integers(C.OFF_B2, C.OFF_W3, 2),
                          rng.integers(C.OFF_W3, C.OFF_B3, 3), [C.OFF_B3, C.OFF_B3 + 2]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (C.loss_and_grad(wp, norm, X, Y)[0] - C.loss_and_grad(wm, norm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
    # the loss is the mean squared error of the forward pass in normalised units
    y = C.forward(w, norm, X)
    assert abs(loss - np.mean(((y - Y) / norm[2]) ** 2)) < 1e-12


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = C.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1 + 2 ** -6) and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3  # half a unit in the 8-bit significand


def test_adam_first_step_is_signed_lr():
    w = np.zeros(5)
    g = np.array([1e-3, -2.0, 5.0, -1e-2, 0.3])
    w1, m, v = C.adam(w, g, np.zeros(5), np.zeros(5), 1, 0.01)
    assert np.allclose(w1, -0.01 * np.sign(g), rtol=1e-4)


def test_synthesis_then_projection_is_identity():
    """Modes -> point values -> modes recovers band-limited residuals (pins the basis norms,
    the R2C scaling and the mask)."""
    M, N, L = 32, 32, 6
    st = random_state(M, N, L, seed=3, amp=1.0)
    R = [st[0], st[1], st[2]]
    z, wz = O.gauss_legendre(4 * L + 16)
    r = C.residual_points(R, M, N, z)
    back = C.project_points(r, M, N, L, z, wz)
    for a, b in zip(back, R):
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()


def test_features_derivatives():
    """The derivative features are the derivatives of the value features: d_z theta by finite
    differences in z; d_x and d_y of theta, u, v, w by shifting the Fourier series of the whole
    state in x or y; lap_h theta and lap_h w as d_x (d_x .) + d_y (d_y .) of the first-derivative
    features; the last column is z."""
    M, N, L, Lx, Ly = 32, 32, 6, 4.0, 3.0
    st = random_state(M, N, L, seed=4, amp=0.1)
    z = np.array([-0.2, 0.05, 0.3])
    h = 1e-6
    F = C.features(st, M, N, Lx, Ly, z)
    assert F.shape == (M * N * len(z), C.NIN)
    Fp = C.features(st, M, N, Lx, Ly, z + h)
    Fm = C.features(st, M, N, Lx, Ly, z - h)
    dz = (Fp[:, 3] - Fm[:, 3]) / (2 * h)
    assert np.abs(dz - F[:, 6]).max() < 1e-6 * np.abs(F[:, 6]).max()
    kx = np.arange(M // 2 + 1) * 2 * np.pi / Lx
    ky = np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N) * 2 * np.pi / Ly
    d = 1e-5

    def shifted(k, axis, s):  # the state moved by s along x (axis 2) or y (axis 1)
        shape = [1, 1, 1]
        shape[axis] = -1
        return [a * np.exp(1j * k * s).reshape(shape) for a in st[:3]] + list(st[3:])
    # central differences: O(d^2) truncation
    Fx = (C.features(shifted(kx, 2, d), M, N, Lx, Ly, z) - C.features(shifted(kx, 2, -d), M, N, Lx, Ly, z)) / (2 * d)
    Fy = (C.features(shifted(ky, 1, d), M, N, Lx, Ly, z) - C.features(shifted(ky, 1, -d), M, N, Lx, Ly, z)) / (2 * d)

    def close(a, b, tol=1e-6):
        return np.abs(a - b).max() < tol * np.abs(b).max()
    # value column -> (d_x column, d_y column)
    for val, cx, cy in ((3, 4, 5), (0, 7, 8), (1, 9, 10), (2, 11, 12)):
        assert close(Fx[:, val], F[:, cx]) and close(Fy[:, val], F[:, cy]), val
    assert close(Fx[:, 4] + Fy[:, 5], F[:, 13])     # lap_h theta
    assert close(Fx[:, 11] + Fy[:, 12], F[:, 14])   # lap_h w
    assert np.all(F[:, 15] == np.repeat(z, M * N))


def test_truncation_keeps_shared_modes():
    st = random_state(64, 64, 8, seed=5)
    t = C.truncate(st[0], 32, 32, 4)
    keep = C.kept_mask(32, 32)
    assert np.all(t[:, ~keep] == 0)
    assert t[2, 3, 5] == st[0][2, 3, 5] and t[1, 32 - 4, 2] == st[0][1, 64 - 4, 2]
