"""CPU tests pinning the FP64 oracle (oracle/rbc_oracle.c) against the reference's own
tables, textbook results and an independent restatement of the nonlinear term."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import (collocation_nonlinear, hermitian_fix, kept_mask, random_state,
                           rel_err)

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_DUMP = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "ref_coeff_dump")


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


@pytest.mark.parametrize("L", [2, 4, 8])
def test_tables_match_reference_coeff(L):
    """Oracle tables vs the reference's own Coeff<double> (coeff.cu compiled unmodified)."""
    g = np.load(os.path.join(GOLD, f"coeff_ref_L{L}.npz"))
    assert np.array_equal(np.array([O.lam(k + 1) for k in range(L)]), g["lambda"])
    t2, t3 = O.coeff_tables(L)
    # the linear operators' tables are restated in closed form: bit-identical
    assert np.array_equal(t2["Iac"], g["Iac"])
    assert np.array_equal(t2["Ia1"], g["Ia1"])
    for n in ("Iab", "Ia2"):
        assert rel_err(t2[n], g[n]) < 1e-12
    for n in O.THREE_NAMES:
        assert rel_err(t3[n], g[n]) < 1e-13, n


def test_iac_quirk_even_even_offdiagonal_zero():
    """coeff.cu:112 computes (lambdacotj - lambdacotj): even-even off-diagonals are 0."""
    t2, _ = O.coeff_tables(8)
    iac = t2["Iac"]
    for i in range(1, 8, 2):
        for j in range(1, 8, 2):
            if i != j:
                assert iac[i, j] == 0.0
    # the odd-odd off-diagonals are not zero
    assert abs(iac[0, 2]) > 1.0


def test_inline_known_constants():
    """Constants quoted at linear_implicit_operator.cu:3-9."""
    t2, _ = O.coeff_tables(2)
    assert abs(t2["Ia1"][0, 0] - 0.6973) < 1e-4
    assert abs(t2["Ia1"][1, 1] + 0.6904) < 1e-4
    assert abs(t2["Iac"][0, 0] + 12.30) < 1e-2
    assert abs(t2["Iac"][1, 1] + 46.05) < 1e-2
    assert abs(O.lam(1) - 4.73) < 1e-2 and abs(O.lam(2) - 7.85) < 1e-2
    assert O.lam(9) == np.pi * 9.5


@pytest.mark.skipif(not os.path.exists(REF_DUMP), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("L", [16, 32])
def test_tables_match_reference_build_large(L, tmp_path):
    import subprocess
    from tests.golden.make_golden import read_ref_dump
    out = tmp_path / "c.bin"
    subprocess.check_call([REF_DUMP, str(L), str(out)])
    g = read_ref_dump(str(out))
    t2, t3 = O.coeff_tables(L)
    assert np.array_equal(t2["Iac"], g["Iac"]) and np.array_equal(t2["Ia1"], g["Ia1"])
    for n in O.THREE_NAMES:
        assert rel_err(t3[n], g[n]) < 5e-13, n


def test_three_factor_tables_are_integrals():
    """Tables reproduce brute-force trapezoid integration of the basis products."""
    from tests.helpers import basis_at
    L = 4
    z = np.linspace(-0.5, 0.5, 200001)
    (C, Cp, Cpp), (S, Sp) = basis_at(L, z)
    _, t3 = O.coeff_tables(L)
    trap = lambda f: np.trapezoid(f, z)
    for (i, j, k) in [(0, 0, 1), (1, 2, 3), (2, 1, 0), (3, 3, 3)]:
        assert abs(t3["Ibb1"][i, j, k] - trap(Cp[i] * Cp[j] * S[k])) < 1e-6 * max(1, abs(t3["Ibb1"][i, j, k]))
        assert abs(t3["Iaab"][i, j, k] - trap(C[i] * C[j] * Cp[k])) < 1e-6 * max(1, abs(t3["Iaab"][i, j, k]))
        assert abs(t3["I111"][i, j, k] - trap(S[i] * S[j] * S[k])) < 1e-8


@pytest.mark.parametrize("shape", [(8, 16), (32, 32), (16, 64)])
def test_fft_matches_numpy_and_roundtrips(shape):
    rng = np.random.default_rng(1)
    N, M = shape
    x = rng.standard_normal((N, M))
    s = O.fft_r2c(x)
    assert np.abs(s - np.fft.rfft2(x)).max() < 1e-12 * np.abs(s).max()
    assert np.abs(O.fft_c2r(s, M) / (M * N) - x).max() < 1e-14 * 10


def test_fft_brute_force_dft():
    rng = np.random.default_rng(2)
    N, M = 4, 8
    x = rng.standard_normal((N, M))
    s = O.fft_r2c(x)
    for ky in range(N):
        for kx in range(M // 2 + 1):
            ref = sum(x[y, xx] * np.exp(-2j * np.pi * (kx * xx / M + ky * y / N)) for y in range(N) for xx in range(M))
            assert abs(s[ky, kx] - ref) < 1e-12


@pytest.mark.parametrize("M,N,nx,ny", [(32, 32, 12, 21), (128, 128, 44, 85), (512, 512, 172, 341)])
def test_dealias_mask_counts(M, N, nx, ny):
    """Kept modes of the reference mask (note kx <= 171 at M=512, not 170)."""
    a = np.ones((1, N, M // 2 + 1), np.complex128)
    O.dealias(a)
    kept = np.abs(a[0]) > 0
    assert kept.sum() == nx * ny
    assert kept[0].sum() == nx
    assert kept[:, 0].sum() == ny
    assert np.array_equal(kept, kept_mask(M, N))


def test_make_real_semantics():
    rng = np.random.default_rng(3)
    L, N, mox = 2, 16, 9
    a = rng.standard_normal((L, N, mox)) + 1j * rng.standard_normal((L, N, mox))
    b = a.copy()
    O.make_real(b)
    assert np.all(b[:, 0, 0].imag == 0)
    assert np.all(b[:, N // 2, 0].imag == 0)
    for r in range(1, N // 2):
        assert np.array_equal(b[:, r, 0], np.conj(a[:, N - r, 0]))
    assert np.array_equal(b[:, :, 1:], a[:, :, 1:])
    assert np.array_equal(b[:, N // 2 + 1:, 0], a[:, N // 2 + 1:, 0])


@pytest.mark.parametrize("M,N,L,nq", [(32, 32, 8, 48), (16, 32, 4, 32), (32, 16, 2, 24)])
def test_nonlinear_pairwise_equals_collocation(M, N, L, nq):
    """The oracle's O(L^2) pairwise form (the reference's algorithm) equals an independent
    collocation form (level fields, pointwise products, quadrature projection)."""
    st = random_state(M, N, L, seed=1, amp=1e-2)
    S = O.Stepper(M, N, L, 4.032, 3.5, 5000.0, 0.7, 1e-3, nthreads=2)
    out = S.nonlinear(*st)
    ref = collocation_nonlinear(*st, M, N, 4.032, 3.5, 0.7, nq)
    keep = kept_mask(M, N)
    for i in range(3):
        assert rel_err(out[i][:, keep], ref[i][:, keep]) < 1e-13, i
    for i in (3, 4):
        assert rel_err(out[i], ref[i]) < 1e-13, i


def test_nonlinear_is_quadratic():
    M, N, L = 16, 16, 4
    st = random_state(M, N, L, seed=5, amp=1e-2)
    S = O.Stepper(M, N, L, 3.0, 3.0, 2000.0, 1.0, 1e-3)
    a = S.nonlinear(*st)
    b = S.nonlinear(*[3.0 * x for x in st])
    for i in range(5):
        assert rel_err(b[i], 9.0 * a[i]) < 1e-13


def _growth_rate(Ra, kc=3.117, L=8):
    Lx = 2 * np.pi / kc
    S = O.Stepper(8, 8, L, Lx, Lx, Ra, 1.0, 1e-4)
    blk = S.implicit_factor(1, 0)  # kx = 1 -> |k| = kc
    Cinv = np.block([[blk[0], blk[1]], [blk[2], blk[3]]])
    mu = np.linalg.eigvals(Cinv)
    sigma = (1.0 - 1.0 / mu) / 1e-4
    return sigma.real.max()


def test_linear_onset_rigid_rigid():
    """Rigid-rigid Rayleigh-Benard onset: Ra_c = 1707.76 at k_c = 3.117 (Chandrasekhar 1961).
    Growth rates come from eigenvalues of the oracle's implicit factors (I - dt B^-1 L)^-1."""
    assert _growth_rate(1690.0) < 0.0 < _growth_rate(1725.0)
    lo, hi = 1690.0, 1725.0
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        if _growth_rate(mid) > 0:
            hi = mid
        else:
            lo = mid
    assert abs(0.5 * (lo + hi) - 1707.76) < 0.5


def test_single_mode_linear_decay():
    """Tiny single g and theta modes (Ra = 0) decay exactly by the implicit Euler factors."""
    M, N, L = 16, 16, 4
    Lx, Ly, Pr, dt = 3.0, 2.5, 0.7, 2e-3
    S = O.Stepper(M, N, L, Lx, Ly, 0.0, Pr, dt)
    th, f, g, F, G = [np.zeros((L, N, M // 2 + 1), np.complex128) for _ in range(3)] + [np.zeros(L, complex)] * 2
    g[1, 2, 3] = 1e-9 + 2e-9j
    th[2, 1, 2] = 3e-9
    S.set_state(th, f, g, F, G)
    n = 20
    S.step(n)
    th2, f2, g2, _, _ = S.get_state()
    k2 = (2 * np.pi * 3 / Lx) ** 2 + (2 * np.pi * 2 / Ly) ** 2
    rg = 1.0 / (1.0 + dt * Pr * (k2 + (2 * np.pi) ** 2))
    assert abs(g2[1, 2, 3] / g[1, 2, 3] - rg ** n) < 1e-9
    k2t = (2 * np.pi * 2 / Lx) ** 2 + (2 * np.pi * 1 / Ly) ** 2
    rt = 1.0 / (1.0 + dt * (k2t + (3 * np.pi) ** 2))
    assert abs(th2[2, 1, 2] / th[2, 1, 2] - rt ** n) < 1e-9


def test_step_preserves_invariants():
    from sdc_amd.initial import low_mode_theta
    M, N, L = 32, 32, 8
    st = low_mode_theta(M, N, L, seed=1, amp=1e-1)
    S = O.Stepper(M, N, L, 4.032, 4.032, 5000.0, 1.0, 1e-3, nthreads=2)
    S.set_state(*st)
    S.step(20)
    th, f, g, F, G = S.get_state()
    keep = kept_mask(M, N)
    for a in (th, f, g):
        assert np.all(a[:, ~keep] == 0)
        assert np.all(a[:, 0, 0].imag == 0)
        for r in range(1, N // 2):
            assert np.array_equal(a[:, r, 0], np.conj(a[:, N - r, 0]))
    assert np.all(f[:, 0, 0] == 0) and np.all(g[:, 0, 0] == 0)
    assert np.all(F.imag == 0) and np.all(G.imag == 0)
    assert np.abs(th).max() > 0


def test_nusselt_conduction_state():
    S = O.Stepper(16, 16, 4, 3.0, 3.0, 1e4, 1.0, 1e-3)
    assert np.allclose(S.nusselt(), [1.0, 1.0, 1.0], atol=0, rtol=0)


def test_bundled_state_fixture_invariants():
    """The reference's bundled FP32 state satisfies the dealias mask exactly."""
    d = np.load(os.path.join(GOLD, "data_state_kept.npz"))
    assert list(d["theta_dims"]) == [257, 512, 2]
    for n in ("theta", "f", "g"):
        assert float(d[n + "_removed_band_maxabs"]) == 0.0
    assert d["F"].shape == (2,) and d["G"].shape == (2,)


def test_init_test_state_values():
    """init_test (util/init.cu:263-301) at a few entries, against values typed from a table of
    sin and cos (sin 1 = 0.8414709848078965, sin 7 = 0.6569865987187891, cos 3 =
    -0.9899924966004454), and its zero pattern."""
    th, f, g, F, G = O.init_test_state(32, 32, 8)
    assert th[0, 0, 1] == pytest.approx(0.8414709848078965e-6, rel=1e-15)
    assert th[5, 0, 7] == pytest.approx(0.6569865987187891e-6, rel=1e-15)
    assert f[3, 0, 7] == th[3, 0, 7] and F[3] == pytest.approx(-0.9899924966004454e-6, rel=1e-15)
    assert np.all(th[:, 1:, :] == 0) and np.all(th[:, 0, 0] == 0) and np.all(th.imag == 0)
    assert not np.any(g) and not np.any(G)
    th32 = O.init_test_state(32, 32, 8, fp32=True)[0]
    assert th32[0, 0, 1] == np.float64(np.float32(1e-6) * np.sin(np.float32(1)))


def test_penalization_everywhere_closed_forms():
    """With the mask 1 everywhere the Brinkman terms of a PENALIZED build
    (nonlinear...cu:1097-1343) reduce to closed forms in mode space (each h field is the
    mode field itself): N_theta = -theta / (2 eta); N_f,k = -(1/eta) sum_j (k^4 delta_jk -
    k^2 Iac(j,k)) f_j (the f mass operator); N_g = (ky^2 + kx ky) g / (2 eta) (the reference's
    fourth g summand differentiates in y, :1265-1272); N_F = -F / (2 eta), N_G = -G / (2 eta)."""
    M, N, L, Lx, Ly, Pr, eta = 32, 32, 8, 4.0, 3.0, 0.7, 0.05
    st = random_state(M, N, L, seed=3, amp=1e-2)
    o = O.Stepper(M, N, L, Lx, Ly, 2e4, Pr, 1e-3)
    base = o.nonlinear(*st)
    o.set_penalization(eta, "everywhere")
    d = [p - b for p, b in zip(o.nonlinear(*st), base)]
    o.close()
    kx = (np.arange(M // 2 + 1) * 2 * np.pi / Lx)[None, :] * np.ones((N, 1))
    r = np.arange(N)
    ky = (np.where(r > N // 2, r - N, r) * 2 * np.pi / Ly)[:, None] * np.ones((1, M // 2 + 1))
    k2 = kx ** 2 + ky ** 2
    iac = O.coeff_tables(L)[0]["Iac"]
    th, f, g, F, G = st
    want = [-(0.5 / eta) * th,
            -(1 / eta) * (k2[None] ** 2 * f - k2[None] * np.einsum("jk,jyx->kyx", iac, f)),
            (0.5 / eta) * (ky ** 2 + kx * ky)[None] * g]
    keep = kept_mask(M, N)
    for a, b in zip(d[:3], want):
        assert rel_err(a[:, keep], b[:, keep]) < 1e-13
    assert rel_err(d[3], -(0.5 / eta) * F) < 1e-11 and rel_err(d[4], -(0.5 / eta) * G) < 1e-11


def test_penalization_masks_and_linearity():
    """The penalty (the difference the mask terms make to the nonlinear operator) is linear in
    the state for every mask shape (the rest of the operator is quadratic) and not zero."""
    M, N, L, Lx, Ly, Pr, eta = 32, 32, 4, 4.0, 4.0, 1.0, 0.1
    a = random_state(M, N, L, seed=1, amp=1e-2)
    b = random_state(M, N, L, seed=2, amp=1e-2)
    o = O.Stepper(M, N, L, Lx, Ly, 2e4, Pr, 1e-3)

    def pen(st, shape):
        o.set_penalization(0.0, shape)
        n0 = o.nonlinear(*st)
        o.set_penalization(eta, shape)
        return [x - y for x, y in zip(o.nonlinear(*st), n0)]
    keep = kept_mask(M, N)
    for shape in ("rectangle", "circle", "diskarray"):
        pa, pb = pen(a, shape), pen(b, shape)
        pab = pen([x + 2 * y for x, y in zip(a, b)], shape)
        for i in range(3):
            assert rel_err(pab[i][:, keep], pa[i][:, keep] + 2 * pb[i][:, keep]) < 1e-12
        assert max(np.abs(x).max() for x in pa[:3]) > 0
    o.close()


@pytest.mark.parametrize("shape,inside,outside", [
    ("rectangle", [(0, 0), (2, 10), (31, 5), (10, 29)], [(3, 3), (16, 16), (28, 28)]),
    ("circle", [(0, 0), (31, 31), (16, 2)], [(16, 16), (16, 4), (4, 16)]),
    ("diskarray", [(5, 8), (16, 24), (26, 8)], [(0, 0), (16, 16), (10, 8)]),
])
def test_penalization_mask_points(shape, inside, outside):
    """Mask values at hand-picked grid points (col, row) of a 32 x 32 grid: rectangle border of
    3 points; circle: outside radius 13 around (16, 16); disk array: 3 x 2 disk centres at
    x = 5, 15, 26 and y = 8, 24 with squared distance below 5."""
    M = N = 32
    o = O.Stepper(M, N, 2, 4.0, 4.0, 2e4, 1.0, 1e-3)
    # a single Fourier mode theta(kx=0, ky=0, m=0) = 1: the theta penalty is the mask's transform
    th = np.zeros((2, N, M // 2 + 1), np.complex128)
    th[0, 0, 0] = 1.0
    st = [th, np.zeros_like(th), np.zeros_like(th), np.zeros(2, complex), np.zeros(2, complex)]
    o.set_penalization(1.0, shape)
    nt = o.nonlinear(*st)[0][0]
    o.close()
    # theta_0 physical = 1 everywhere, so N_theta,0 = -(1/2) R2C(mask / (M N)) (unnormalised R2C)
    mask = -2.0 * (M * N) * np.fft.irfft2(nt, s=(N, M))
    for col, row in inside:
        assert abs(mask[row, col] - 1.0) < 1e-12, (col, row)
    for col, row in outside:
        assert abs(mask[row, col]) < 1e-12, (col, row)
