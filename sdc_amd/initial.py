"""Seeded initial states in the reference layout.

The reference seeds with cuRAND XORWOW (util/init.cu:9-20), which cannot be reproduced
outside that library; states here come from a counter-based generator instead:
SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
OOPSLA 2014) on the key seed * 2^32 + 2 * (flat index) + {0, 1}, mapped to two uniforms in
(0, 1] and turned into one complex Gaussian by Box-Muller.  The same (seed, shape) always
gives the same state on every host.

Fields are numpy complex128 arrays shaped (L, N, M/2+1) (reference layout, slowest axis
first); F and G are (L,).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    """SplitMix64 output function on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + _GOLDEN).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(30))) * _M1).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(27))) * _M2).astype(np.uint64)
        return z ^ (z >> np.uint64(31))


def gaussian_complex(seed, n):
    """n complex standard normals (E|z|^2 = 2) from (seed, index)."""
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(1 << 32)
        a = splitmix64(base + np.uint64(2) * idx)
        b = splitmix64(base + np.uint64(2) * idx + np.uint64(1))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) / 9007199254740992.0  # (0, 1]
    u2 = (b >> np.uint64(11)).astype(np.float64) / 9007199254740992.0          # [0, 1)
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)


def kept_mask(M, N):
    """The reference dealias mask (implicit_explicit_stepping.cu:163-182): True = kept."""
    mox = M // 2 + 1
    col = np.arange(mox)[None, :]
    row = np.arange(N)[:, None]
    removed = (mox - col <= (1.0 / 3.0) * mox) | (np.abs(0.5 * N - row) <= (1.0 / 3.0) * 0.5 * N)
    return ~removed


def make_real(a):
    """make_real (util.cu) on an (L, N, mox) array, in place: column kx=0 row r <- conj(row N-r)
    for 0 < r < N/2, and imaginary parts of rows 0 and N/2 cleared."""
    N = a.shape[1]
    a[:, 0, 0] = a[:, 0, 0].real
    for r in range(1, N // 2):
        a[:, r, 0] = np.conj(a[:, N - r, 0])
    a[:, N // 2, 0] = a[:, N // 2, 0].real
    return a


def low_mode_theta(M, N, L, seed=1, amp=1e-3, nkx=4, nky=4, nm=2):
    """theta in the lowest nkx x nky x nm modes (kx = 0..nkx-1, ky = 0..nky-1, m = 0..nm-1),
    amplitude amp * N(0,1); at kx = 0 the rows ky > 0 are mirrored to N - ky as conjugates
    so that make_real keeps them.  f = g = F = G = 0."""
    mox = M // 2 + 1
    th = np.zeros((L, N, mox), np.complex128)
    z = gaussian_complex(seed, nm * nky * nkx).reshape(nm, nky, nkx) * amp
    nm_ = min(nm, L)
    th[:nm_, :nky, :nkx] = z[:nm_]
    for r in range(1, nky):
        th[:nm_, N - r, 0] = np.conj(th[:nm_, r, 0])
    th[:, 0, 0] = th[:, 0, 0].real
    th[:, ~kept_mask(M, N)] = 0.0
    make_real(th)
    zeros = np.zeros((L, N, mox), np.complex128)
    return [th, zeros.copy(), zeros.copy(), np.zeros(L, np.complex128), np.zeros(L, np.complex128)]


def random_theta(M, N, L, seed=3, amp=1e-3):
    """theta ~ amp * N(0,1) in every kept mode (Hermitian-consistent), other fields zero."""
    mox = M // 2 + 1
    th = gaussian_complex(seed, L * N * mox).reshape(L, N, mox) * amp
    th[:, ~kept_mask(M, N)] = 0.0
    make_real(th)
    zeros = np.zeros((L, N, mox), np.complex128)
    return [th, zeros.copy(), zeros.copy(), np.zeros(L, np.complex128), np.zeros(L, np.complex128)]
