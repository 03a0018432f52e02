"""CPU pins of oracle/patterns.py (the pattern starts of util/init.cu): the SplitMix64 restatement
against the generator's published outputs, and every pattern against closed forms the
mathematics fixes (a single Fourier bin, the phase of the cores' atan2 terms on the lines through
them), so that a wrong sign, core position, index or normalisation fails."""
import numpy as np
import pytest

from oracle import patterns as PT

AMP = 1e-6


def test_splitmix64_published_outputs():
    # SplitMix64 from state 0: the first outputs of the reference generator
    g = 0x9E3779B97F4A7C15
    got = [int(PT.splitmix64_mix(np.uint64((k * g) % (1 << 64)))) for k in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_range_and_moments():
    u = PT.uniform(7, np.arange(200000, dtype=np.uint64))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 0.01 / 12


@pytest.mark.parametrize("M,N,mode", [(64, 32, 3), (128, 64, 5)])
def test_isr_is_one_fourier_bin(M, N, mode):
    """amp cos(2 pi mode x + phi) has one nonzero R2C bin (ky = 0, kx = mode) equal to
    amp M N / 2 exp(i phi), phi = 2 pi u(seed, 2^64 - 1)."""
    seed = 11
    S = PT.spectrum("isr", M, N, mode, AMP, seed)
    phi = 2 * np.pi * float(PT.uniform(seed, np.uint64(0xFFFFFFFFFFFFFFFF)))
    want = AMP * M * N / 2 * np.exp(1j * phi)
    assert abs(S[0, mode] - want) < 1e-12 * abs(want)
    S[0, mode] = 0
    assert np.abs(S).max() < 1e-12 * abs(want)


def test_rectangle_noise_columns_and_stripes():
    M, N, mode = 64, 32, 4
    P = PT.plane("rectangle", M, N, mode, AMP, seed=3)
    i = np.arange(M)
    assert np.allclose(P[:, M // 8:], AMP * np.cos(2 * np.pi * mode * i[M // 8:] / M)[None, :], rtol=0, atol=1e-18)
    strip = P[:, : M // 8]
    assert np.abs(strip).max() <= AMP and np.unique(strip).size == strip.size  # noise, no repeats
    assert PT.rectangle_mode(4.032, 32) == int((4.032 / 31.0) * 27.0 / 2.0) == 1


def test_giantspiral_on_the_line_through_the_cores():
    """y = 1/2: left of (0.5, 0.5) both atan2 are pi and cancel, between the cores they are 0 and
    pi, right of (0.8, 0.5) both 0: amp cos(2 pi mode r), -amp cos(2 pi mode r), amp cos(...)."""
    mode = 3
    for x, sgn in ((0.2, 1.0), (0.65, -1.0), (0.9, 1.0)):
        v = PT.value("giantspiral", x, 0.5, mode, AMP)
        assert abs(v - sgn * AMP * np.cos(2 * np.pi * mode * abs(x - 0.5))) < 1e-18


def test_sv_phase_at_y0():
    """y = 0: phase = -x^2 left of x_left = 0.4, -x^2 + 1 right of x_right = 0.6 and
    -x^2 + (1 - cos(pi (x - 0.4) / 0.2)) / 2 between."""
    mode = 2
    for x in (0.1, 0.35, 0.45, 0.55, 0.7, 0.95):
        ph = -x * x + (1.0 if x > 0.6 else (0.5 * (1 - np.cos(np.pi * (x - 0.4) / 0.2)) if x > 0.4 else 0.0))
        assert abs(PT.value("sv", x, 0.0, mode, AMP) - AMP * np.cos(2 * np.pi * ph)) < 1e-18


def test_dislocation_cores_on_their_vertical_line():
    """On x = x1 (the line through both cores): between the cores the atan2 terms are +pi/2 and
    -pi/2 (the pattern is -amp cos(2 pi mode x1)), outside they cancel (+amp cos)."""
    M, N, mode = 64, 32, 5
    dx, dy = 1.0 / M, 1.0 / N
    x1 = 0.5 + dx / 2
    y1, y2 = 0.25 + dy / 2, 0.75 + dy / 2
    base = AMP * np.cos(2 * np.pi * mode * x1)
    assert abs(PT.value("dislocation", x1, 0.5 * (y1 + y2), mode, AMP, M, N) + base) < 1e-18
    assert abs(PT.value("dislocation", x1, 0.1, mode, AMP, M, N) - base) < 1e-18
    assert abs(PT.value("dislocation", x1, 0.95, mode, AMP, M, N) - base) < 1e-18


def test_cr_cores_on_their_horizontal_line():
    """On y = y1: between the cores (x1 < x < x2) the atan2 terms are 0 and pi (the pattern is
    -amp cos(2 pi mode y1)), right of both they are 0 (+amp cos)."""
    M, N, mode = 64, 32, 3
    x1 = 0.5 * (0.2 * M + 1) / M
    y1 = 0.5 * (N + 1) / N
    base = AMP * np.cos(2 * np.pi * mode * y1)
    assert abs(PT.value("cr", 0.5, y1, mode, AMP, M, N) + base) < 1e-18
    assert abs(PT.value("cr", 1.0 - x1 / 2, y1, mode, AMP, M, N) - base) < 1e-18


@pytest.mark.parametrize("kind", list(PT.KINDS))
def test_planes_bounded_and_real_spectrum(kind):
    M, N = 64, 32
    P = PT.plane(kind, M, N, 3, AMP, seed=5)
    assert P.shape == (N, M) and np.abs(P).max() <= AMP * (1 + 1e-12)
    S = PT.spectrum(kind, M, N, 3, AMP, seed=5)
    assert np.allclose(np.fft.irfft2(S, s=(N, M)), P, rtol=0, atol=1e-20)
