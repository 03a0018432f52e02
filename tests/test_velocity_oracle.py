"""CPU checks that pin the FP64 restatement of the velocity operator, the tracer schemes and
the energy spectra (oracle/velocity.py) by mathematics independent of its own formulas."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import velocity as V
from tests.helpers import random_state

M, N, L, LX, LY = 32, 32, 6, 4.0, 3.0


@pytest.fixture(scope="module")
def state():
    O.build()
    return random_state(M, N, L, seed=11, amp=1e-1)


def test_velocity_is_divergence_free(state):
    """The poloidal/toroidal form makes div u = 0 for any state: a sign or basis slip in any of
    u_x, u_y, u_z (or the mean flow) breaks it."""
    z = np.linspace(-0.45, 0.45, 7)
    h = 1e-5
    u, v, _ = V.velocity_spectral(state, M, N, LX, LY, z)
    _, _, wp = V.velocity_spectral(state, M, N, LX, LY, z + h)
    _, _, wm = V.velocity_spectral(state, M, N, LX, LY, z - h)
    kx, ky = V.wavenumbers(M, N, LX, LY)
    KX, KY = np.meshgrid(kx, ky)
    div = 1j * KX * u + 1j * KY * v + (wp - wm) / (2 * h)
    scale = np.abs(1j * KX * u).max()
    assert np.abs(div).max() < 1e-7 * scale


def test_no_slip_and_impermeable_plates(state):
    """C_m and C'_m vanish at z = +-1/2 (rigid plates) and so do u_z and the poloidal part of
    u_h; the toroidal part S_m vanishes there too."""
    C, Cp, S = V.basis(L, [-0.5, 0.5])
    assert np.abs(C).max() < 1e-10 and np.abs(Cp).max() < 1e-8 and np.abs(S).max() < 1e-15
    ux, uy, uz = V.velocity_grid(state, M, N, LX, LY, [-0.5, 0.5])
    assert max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max()) < 1e-8


def test_points_equal_grid_at_nodes(state):
    """Direct Fourier sums at grid nodes equal the C2R grid (two independent syntheses)."""
    z = np.array([-0.3, 0.1, 0.42])
    grid = V.velocity_grid(state, M, N, LX, LY, z)
    rng = np.random.default_rng(0)
    iz = rng.integers(0, 3, 40)
    iy = rng.integers(0, N, 40)
    ix = rng.integers(0, M, 40)
    u, v, w = V.velocity_points(state, M, N, LX, LY, ix * LX / M, iy * LY / N, z[iz])
    for got, g in zip((u, v, w), grid):
        ref = g[iz, iy, ix]
        assert np.abs(got - ref).max() < 1e-12 * np.abs(g).max()


def test_points_are_periodic(state):
    x = np.array([0.3, 1.7]); y = np.array([0.2, 2.9]); z = np.array([0.0, -0.2])
    a = V.velocity_points(state, M, N, LX, LY, x, y, z)
    b = V.velocity_points(state, M, N, LX, LY, x + 3 * LX, y - 2 * LY, z)
    for p, q in zip(a, b):
        assert np.allclose(p, q, rtol=1e-10, atol=1e-13)


def _linear_grid(a, b, c, d, xs=16, ys=12):
    zl = V.linspace_levels()
    iz, iy, ix = np.meshgrid(np.arange(V.ZNUM), np.arange(ys), np.arange(xs), indexing="ij")
    f = a + b * ix + c * iy + d * iz
    return [f, 2 * f, -f], zl


def test_trilinear_is_exact_for_linear_fields():
    xs, ys = 16, 12
    grid, zl = _linear_grid(0.3, 0.7, -0.2, 0.05, xs, ys)
    rng = np.random.default_rng(1)
    # interior points (no periodic wrap between the two x / y neighbours)
    x = rng.uniform(0, (xs - 1.0) / xs * LX, 50)
    y = rng.uniform(0, (ys - 1.0) / ys * LY, 50)
    z = rng.uniform(-V.ZMAX, V.ZMAX, 50)
    u, v, w = V.trilinear(grid, LX, LY, x, y, z)
    xg, yg = x * xs / LX, y * ys / LY
    zg = (V.ZMAX + z) / (2 * V.ZMAX) * (V.ZNUM - 1)
    ref = 0.3 + 0.7 * xg - 0.2 * yg + 0.05 * zg
    assert np.abs(u - ref).max() < 1e-12 and np.abs(v - 2 * ref).max() < 1e-12 and np.abs(w + ref).max() < 1e-12
    # periodic lookup: a shift by whole boxes changes nothing
    u2, _, _ = V.trilinear(grid, LX, LY, x - 2 * LX, y + LY, z)
    assert np.abs(u2 - u).max() < 1e-12


def test_rk4_matches_the_exponential_to_fourth_order():
    """u = v = 0, w = c z (linear in z, reproduced exactly by the trilinear lookup): one RK4 step
    multiplies z by 1 + h + h^2/2 + h^3/6 + h^4/24, h = c dt."""
    xs, ys = 8, 8
    zl = -V.ZMAX + np.arange(V.ZNUM) * (2 * V.ZMAX / (V.ZNUM - 1))  # the lookup's own z map
    c = 3.0
    w = np.broadcast_to((c * zl)[:, None, None], (V.ZNUM, ys, xs)).copy()
    grid = [np.zeros_like(w), np.zeros_like(w), w]
    z0 = np.array([-0.2, 0.01, 0.1])
    x0 = np.array([0.1, 1.0, 2.0]); y0 = np.array([0.5, 1.5, 2.5])
    dt = 0.05
    x1, y1, z1 = V.rk4_step(grid, LX, LY, x0, y0, z0, dt)
    h = c * dt
    assert np.abs(z1 - z0 * (1 + h + h * h / 2 + h ** 3 / 6 + h ** 4 / 24)).max() < 1e-13
    assert np.array_equal(x1, x0) and np.array_equal(y1, y0)


def test_lid_reflection():
    z = np.array([0.2, V.ZMAX + 0.01, 2 * V.ZMAX + 0.1, -V.ZMAX - 0.02, -3.0, np.nan])
    r = V.reflect_lids(z)
    assert r[0] == 0.2
    assert abs(r[1] - (V.ZMAX - 0.01)) < 1e-15
    assert r[2] == V.ZMAX
    assert abs(r[3] - (-V.ZMAX + 0.02)) < 1e-15
    assert r[4] == -V.ZMAX
    assert r[5] == V.ZMAX


def test_energy_totals_equal_physical_space_mean(state):
    """Parseval: the shell sums equal (1/2) <u^2 + v^2> over the grid, integrated over z with
    the same Gauss-Legendre rule (velocity synthesised by the C2R path)."""
    nq = 24
    E, (eh, ez) = V.energy_spectrum(state, M, N, LX, LY, nq)
    zq, wq = O.gauss_legendre(nq)
    ux, uy, uz = V.velocity_grid(state, M, N, LX, LY, zq)
    eh_phys = 0.5 * np.sum(wq * (ux ** 2 + uy ** 2).mean(axis=(1, 2)))
    ez_phys = 0.5 * np.sum(wq * (uz ** 2).mean(axis=(1, 2)))
    assert abs(E.sum() - eh) < 1e-14 * eh
    assert abs(eh - eh_phys) < 1e-12 * eh_phys
    assert abs(ez - ez_phys) < 1e-12 * ez_phys


def test_reference_binning_single_mode():
    """One poloidal mode (kx index 2, ky index -3) lands in bin floor(|k| / inc) with
    |k_max| / size = inc; every component's energy sits in that one bin."""
    st = [np.zeros((L, N, M // 2 + 1), complex) for _ in range(3)] + [np.zeros(L, complex)] * 2
    st[1][0, N - 3, 2] = 0.4 + 0.1j
    E3 = V.energy_spectrum_ref(st, M, N, LX, LY, 0.1)
    size = max(M // 2 + 1, N // 2 + 1)
    inc = np.hypot(2 * np.pi * (M // 2) / LX, 2 * np.pi * (N // 2) / LY) / size
    b = int(np.hypot(2 * np.pi * 2 / LX, 2 * np.pi * 3 / LY) // inc)
    nzb = np.nonzero(E3.sum(axis=0))[0]
    assert list(nzb) == [b]
    u, v, w = V.velocity_spectral(st, M, N, LX, LY, [0.1])
    assert abs(E3[0, b] - np.abs(u[0, N - 3, 2]) ** 2) < 1e-15
    assert abs(E3[2, b] - np.abs(w[0, N - 3, 2]) ** 2) < 1e-15


def test_state_file_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3, 8, 5)) + 1j * rng.standard_normal((3, 8, 5))
    for dt in (np.complex64, np.complex128):
        p = tmp_path / f"a_{np.dtype(dt).itemsize}.bin"
        V.write_state_file(p, a, dt)
        dims, vals = V.read_state_file(p)
        assert dims == [5, 8, 3]
        assert np.array_equal(vals.reshape(3, 8, 5), a.astype(dt).astype(np.complex128))
        assert p.stat().st_size == 20 + a.size * np.dtype(dt).itemsize


def test_bundled_state_header_layout():
    """The reference's own files (gpu_flow/data/*_00030000.bin, kept as a fixture) carry
    dims {M/2+1, N, L} = {257, 512, 2}: the writer's dims order."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "data_state_kept.npz"))
    assert d["theta_dims"].tolist() == [257, 512, 2]
    assert d["F_dims"].tolist() == [1, 1, 2]


def test_plane_operators_special_masks():
    """oracle/planes.py pinned by what the operators do under uniform masks: a mask of ones
    leaves a band-limited state unchanged; with zeros the freeze gives the frozen state, the heat
    source adds 0.015 to every mode's mean (the transform of a constant 0.015 over M N points,
    scaled by 1/(M N)), and the scale multiplies by the factor."""
    from oracle import planes as PL
    from tests.helpers import random_state
    M, N, L = 32, 16, 3
    st = random_state(M, N, L, seed=4, amp=1e-2)
    other = random_state(M, N, L, seed=5, amp=1e-2)
    ones = np.ones((N, M), np.int32)
    zeros = np.zeros((N, M), np.int32)
    for a, b in zip(PL.apply(st[:3], "heat", ones, M, N), st[:3]):
        assert np.abs(a - b).max() < 1e-15
    frozen = PL.capture(other[:3], M, N)
    for a, b in zip(PL.apply(st[:3], "freeze", zeros, M, N, frozen=frozen), other[:3]):
        assert np.abs(a - b).max() < 1e-15
    heated = PL.apply(st[:1], "heat", zeros, M, N)[0]
    d = heated - st[0]
    assert np.allclose(d[:, 0, 0], 0.015, atol=1e-15) and np.abs(d[:, 1:, :]).max() < 1e-15
    scaled = PL.apply(st[:1], "scale", zeros, M, N, factor=1.1)[0]
    assert np.abs(scaled - 1.1 * st[0]).max() < 1e-15


def test_plane_masks():
    """Mask geometry: circle of radius 0.8 * min/2 around the centre (freeze), its complement
    kept; the rectangle of width 0.1 keeps a border of ceil(0.1 M) points (masktemp)."""
    from oracle import planes as PL
    m = PL.plane_mask("circle", 64, 32, 0.8)
    assert m[16, 32] == 0 and m[0, 0] == 1 and m[16, 32 + 13] == 1 and m[16, 32 + 12] == 0
    r = PL.plane_mask("rectangle", 64, 32, 0.1)
    assert r[16, 6] == 1 and r[16, 7] == 0 and r[3, 32] == 1 and r[4, 32] == 0 and r[16, 58] == 1
    t = PL.plane_mask("circle", 64, 32, 0.2, 0.2, 0.3)
    assert t[int(0.3 * 32), int(0.2 * 64)] == 0 and t[31, 63] == 1
