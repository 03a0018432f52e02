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
    """d_z theta and d_x theta features are the derivatives of the theta feature (finite
    differences in z, and a shift of the Fourier series in x)."""
    M, N, L, Lx, Ly = 32, 32, 6, 4.0, 3.0
    st = random_state(M, N, L, seed=4, amp=0.1)
    z = np.array([-0.2, 0.05, 0.3])
    h = 1e-6
    F = C.features(st, M, N, Lx, Ly, z)
    Fp = C.features(st, M, N, Lx, Ly, z + h)
    Fm = C.features(st, M, N, Lx, Ly, z - h)
    dz = (Fp[:, 3] - Fm[:, 3]) / (2 * h)
    assert np.abs(dz - F[:, 6]).max() < 1e-6 * np.abs(F[:, 6]).max()
    kx = np.arange(M // 2 + 1) * 2 * np.pi / Lx
    dx = 1e-6
    shifted = [st[0] * np.exp(1j * kx * dx)[None, None, :]] + list(st[1:])
    Fs = C.features(shifted, M, N, Lx, Ly, z)
    assert np.abs((Fs[:, 3] - F[:, 3]) / dx - F[:, 4]).max() < 1e-5 * np.abs(F[:, 4]).max()
    assert np.all(F[:, 7] == np.repeat(z, M * N))


def test_truncation_keeps_shared_modes():
    st = random_state(64, 64, 8, seed=5)
    t = C.truncate(st[0], 32, 32, 4)
    keep = C.kept_mask(32, 32)
    assert np.all(t[:, ~keep] == 0)
    assert t[2, 3, 5] == st[0][2, 3, 5] and t[1, 32 - 4, 2] == st[0][1, 64 - 4, 2]
