"""FP64 numpy restatement of the learned closure (sdc_amd/csrc/rbc_closure.hip): the network's
forward pass, loss gradient and Adam update, the BF16 rounding of its weights, and the
synthesis / projection that move between coarse modes and coarse points.

TEST INFRASTRUCTURE: only tests/ imports this module.

Weight vector layout (4931 floats): W1[64][8], b1[64], W2[64][64], b2[64], W3[3][64], b3[3];
hidden activations tanh; loss = mean over points and outputs of (y - t / out_scale)^2 with
inputs normalised as (x - in_mean) * in_inv and predictions y * out_scale.
"""
import numpy as np

from oracle import oracle as O
from oracle import velocity as V

NIN, NH, NOUT = 8, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NW = OFF_B3 + NOUT


def bf16_round(w):
    """float32 -> nearest BF16 (ties to even) -> float32."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(w):
    w = np.asarray(w, np.float64)
    return (w[OFF_W1:OFF_B1].reshape(NH, NIN), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(NH, NH),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3].reshape(NOUT, NH), w[OFF_B3:NW])


def forward(w, norm, X):
    """Predictions in physical units, (n, 3)."""
    mean, inv, scale = (np.asarray(a, np.float64) for a in norm)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    x = (np.asarray(X, np.float64) - mean) * inv
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return (h2 @ W3.T + b3) * scale


def loss_and_grad(w, norm, X, Y):
    """Normalised mean squared error and its gradient with respect to the flat weights."""
    mean, inv, scale = (np.asarray(a, np.float64) for a in norm)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    x = (np.asarray(X, np.float64) - mean) * inv
    t = np.asarray(Y, np.float64) / scale
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    n = len(x)
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / (n * NOUT)
    g = np.zeros(NW)
    g[OFF_W3:OFF_B3] = (dy.T @ h2).ravel()
    g[OFF_B3:NW] = dy.sum(axis=0)
    d2 = (dy @ W3) * (1 - h2 * h2)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(axis=0)
    d1 = (d2 @ W2) * (1 - h1 * h1)
    g[OFF_W1:OFF_B1] = (d1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(axis=0)
    return loss, g


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


# ---- coarse modes <-> coarse points -------------------------------------------------------

def kept_mask(M, N):
    """The dealias mask (implicit_explicit_stepping.cu:163-182): True where a mode is kept."""
    mox = M // 2 + 1
    col = np.arange(mox)[None, :]
    row = np.arange(N)[:, None]
    return ~((mox - col <= mox / 3.0) | (np.abs(0.5 * N - row) <= N / 6.0))


def truncate(field, Mc, Nc, Lc):
    """P_c: a fine (Lf, Nf, Mf/2+1) field (or an (Lf,) vector) to the coarse kept modes."""
    a = np.asarray(field)
    if a.ndim == 1:
        return a[:Lc].copy()
    Lf, Nf, moxf = a.shape
    out = np.zeros((Lc, Nc, Mc // 2 + 1), complex)
    keep = kept_mask(Mc, Nc)
    for row in range(Nc):
        ky = row if row <= Nc // 2 else row - Nc
        frow = ky % Nf
        for col in range(Mc // 2 + 1):
            if keep[row, col]:
                out[:, row, col] = a[:Lc, frow, col]
    return out


def _synth(coefs, M, N):
    """Unnormalised C2R of (nz, N, mox) spectra: (nz, N, M) physical values."""
    return np.fft.irfft2(coefs, s=(N, M)) * (M * N)


def features(state, M, N, Lx, Ly, z):
    """Inputs at the points (x, y, z_l), row-major pt = (l N + y) M + x: (npts, 8) =
    u, v, w, theta, d_x theta, d_y theta, d_z theta, z."""
    th = state[0]
    L = th.shape[0]
    u, v, w = V.velocity_grid(state, M, N, Lx, Ly, z)
    kx, ky = V.wavenumbers(M, N, Lx, Ly)
    KX, KY = np.meshgrid(kx, ky)
    C, Cp, S = V.basis(L, z)
    Sp = np.zeros_like(S)
    for i in range(L):
        Sp[i] = np.pi * (i + 1) * np.cos(np.pi * (i + 1) * (np.clip(z, -0.5, 0.5) + 0.5))
    ts = np.einsum("mz,myx->zyx", S, th)
    tz = np.einsum("mz,myx->zyx", Sp, th)
    fields = [u, v, w, _synth(ts, M, N), _synth(1j * KX * ts, M, N), _synth(1j * KY * ts, M, N), _synth(tz, M, N)]
    nz = len(z)
    zz = np.broadcast_to(np.asarray(z, np.float64)[:, None, None], (nz, N, M))
    return np.stack([f.ravel() for f in fields] + [zz.ravel()], axis=1)


def residual_points(R, M, N, z):
    """r_theta = sum S_m R_theta,m, r_f = sum C_m R_f,m, r_g = sum S_m R_g,m at the points."""
    L = R[0].shape[0]
    C, Cp, S = V.basis(L, z)
    cols = [np.einsum("mz,myx->zyx", S, R[0]), np.einsum("mz,myx->zyx", C, R[1]), np.einsum("mz,myx->zyx", S, R[2])]
    return np.stack([_synth(c, M, N).ravel() for c in cols], axis=1)


def project_points(r, M, N, L, z, wz):
    """Modes (theta, f, g) of point values r (npts, 3): R2C / (M N) per level, then
    <., psi_m> / <psi_m, psi_m> by the quadrature (z, wz); psi = S, C, S."""
    nz = len(z)
    C, Cp, S = V.basis(L, z)
    zf, wf = O.gauss_legendre(4 * L + 64)
    Cf, _, Sf = V.basis(L, zf)
    nC = (wf * Cf * Cf).sum(axis=1)
    nS = (wf * Sf * Sf).sum(axis=1)
    out = []
    for f, (B, nrm) in enumerate(((S, nS), (C, nC), (S, nS))):
        vals = r[:, f].reshape(nz, N, M)
        spec = np.fft.rfft2(vals) / (M * N)                       # (nz, N, mox)
        out.append(np.einsum("mz,z,zyx->myx", B / nrm[:, None], np.asarray(wz), spec) * kept_mask(M, N))
    return out
