"""Shared test helpers: seeded spectral states and an independent numpy restatement of the
nonlinear term in collocation form (test infrastructure)."""
import numpy as np


def kept_mask(M, N):
    """Reference dealias mask (implicit_explicit_stepping.cu:163-182): True = kept."""
    mox = M // 2 + 1
    col = np.arange(mox)[None, :]
    row = np.arange(N)[:, None]
    zero = (mox - col <= (1.0 / 3.0) * mox) | (np.abs(0.5 * N - row) <= (1.0 / 3.0) * 0.5 * N)
    return ~zero


def hermitian_fix(a):
    """make_real restated (util.cu make_real) on an (L, N, mox) array, in place."""
    N = a.shape[1]
    a[:, 0, 0] = a[:, 0, 0].real
    for r in range(1, N // 2):
        a[:, r, 0] = np.conj(a[:, N - r, 0])
    a[:, N // 2, 0] = a[:, N // 2, 0].real
    return a


def random_state(M, N, L, seed=0, amp=1e-2, decay=0.0, zero_mean_fg=True):
    """Dealiased, Hermitian-consistent random state (theta, f, g, F, G)."""
    rng = np.random.default_rng(seed)
    mox = M // 2 + 1
    keep = kept_mask(M, N)
    out = []
    kx = np.arange(mox)[None, :]
    ky = np.fft.fftfreq(N, 1.0 / N)[:, None]
    damp = np.exp(-decay * np.sqrt(kx ** 2 + ky ** 2))
    for i in range(3):
        a = (rng.standard_normal((L, N, mox)) + 1j * rng.standard_normal((L, N, mox))) * amp
        a *= damp[None]
        a *= (1.0 / (1.0 + np.arange(L)) ** 2)[:, None, None]
        a[:, ~keep] = 0.0
        hermitian_fix(a)
        if zero_mean_fg and i > 0:
            a[:, 0, 0] = 0.0
        out.append(a)
    F = rng.standard_normal(L) * amp + 0j
    G = rng.standard_normal(L) * amp + 0j
    return out + [F, G]


def _lam(k):
    r = [0.0, 1.505618731142, 2.499752670074, 3.500010679436, 4.499999538484, 5.500000019944,
         6.499999999138, 7.500000000037, 8.499999999998]
    return np.pi * (k + 0.5) if k > 8 else np.pi * r[k]


def basis_at(L, z):
    """(C, C', C'', S, S') at nodes z, each (L, nz)."""
    C = np.zeros((3, L, len(z)))
    S = np.zeros((2, L, len(z)))
    for m in range(L):
        k = m + 1
        lam = _lam(k)
        if k % 2:
            ch, c = np.cosh(lam / 2), np.cos(lam / 2)
            C[0, m] = np.cosh(lam * z) / ch - np.cos(lam * z) / c
            C[1, m] = lam * (np.sinh(lam * z) / ch + np.sin(lam * z) / c)
            C[2, m] = lam ** 2 * (np.cosh(lam * z) / ch + np.cos(lam * z) / c)
        else:
            sh, s = np.sinh(lam / 2), np.sin(lam / 2)
            C[0, m] = np.sinh(lam * z) / sh - np.sin(lam * z) / s
            C[1, m] = lam * (np.cosh(lam * z) / sh - np.cos(lam * z) / s)
            C[2, m] = lam ** 2 * (np.sinh(lam * z) / sh + np.sin(lam * z) / s)
        S[0, m] = np.sin(k * np.pi * (z + 0.5))
        S[1, m] = k * np.pi * np.cos(k * np.pi * (z + 0.5))
    return C, S


def wavenumbers(M, N, Lx, Ly):
    mox = M // 2 + 1
    kx = (np.arange(mox) * (2 * np.pi / Lx))[None, :] * np.ones((N, 1))
    r = np.arange(N)
    r = np.where(r > N // 2, r - N, r)
    ky = (r * (2 * np.pi / Ly))[:, None] * np.ones((1, mox))
    return kx, ky


def collocation_nonlinear(th, f, g, F, G, M, N, Lx, Ly, Pr, nq):
    """Nonlinear term of nonlinear_operator_rayleigh_noslip.cu evaluated in collocation form:
    level fields u, v, w, grad(theta) at nq Gauss-Legendre levels, pointwise products on the
    M x N grid, projection with quadrature weights.  Exact whenever the nq-point rule
    integrates the Galerkin triple products exactly."""
    L = th.shape[0]
    x, w = np.polynomial.legendre.leggauss(nq)
    z, w = x / 2, w / 2
    (Cv, Cp, Cpp), (Sv, Sp) = basis_at(L, z)
    kx, ky = wavenumbers(M, N, Lx, Ly)
    k2 = kx ** 2 + ky ** 2
    ev = lambda B, a: np.einsum("ml,mqp->lqp", B, a)
    Fcp, Fc, Gs, Ts, Tz = ev(Cp, f), ev(Cv, f), ev(Sv, g), ev(Sv, th), ev(Sp, th)
    u = 1j * kx * Fcp + 1j * ky * Gs
    v = 1j * ky * Fcp - 1j * kx * Gs
    u[:, 0, 0] += Sv.T @ F
    v[:, 0, 0] += Sv.T @ G
    wv = k2 * Fc
    tx, ty = 1j * kx * Ts, 1j * ky * Ts
    c2r = lambda a: np.fft.irfft2(a, s=(N, M)) * (M * N)
    r2c = lambda a: np.fft.rfft2(a / (M * N))
    U, V, W, TX, TY, TZ = map(c2r, (u, v, wv, tx, ty, Tz))
    Puu, Puv, Pvv, Puw, Pvw, Pww = (r2c(a) for a in (U * U, U * V, V * V, U * W, V * W, W * W))
    Padv = r2c(U * TX + V * TY + W * TZ)
    proj = lambda B, a: np.einsum("kl,l,lqp->kqp", B, w, a)
    A = kx ** 2 * Puu + 2 * kx * ky * Puv + ky ** 2 * Pvv - k2 * Pww
    Bf = 1j * kx * Puw + 1j * ky * Pvw
    Nf = proj(Cp, A) + proj(Cpp, Bf) + k2 * proj(Cv, Bf)
    Ng = proj(Sv, -kx * ky * (Puu - Pvv) + (kx ** 2 - ky ** 2) * Puv) + proj(Sp, 1j * kx * Pvw - 1j * ky * Puw)
    Nt = proj(Sv, Padv)
    NF = -(Sp * w) @ Puw[:, 0, 0]
    NG = -(Sp * w) @ Pvw[:, 0, 0]
    return [-Nt, -Nf / Pr, -Ng / Pr, -NF / Pr, -NG / Pr]


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())
