"""FP64 numpy restatement of the reference's velocity operator, particle tracer and energy
spectrum, plus the spectral tracer and the z-integrated spectrum of the GPU library.

TEST INFRASTRUCTURE: only tests/ and __graft_entry__.smoke() import this module.

Fields are complex arrays shaped (L, N, M/2+1) (the reference layout read slowest axis
first); F and G are (L,).  Positions are float arrays.
"""
import numpy as np

from oracle import oracle as O

ZNUM = 128                 # PARTICLE_Z_NUM (cuda_defines.h:44)
ZMAX = 0.5 - 1e-3          # PARTICLE_Z_MAX (cuda_defines.h:45), evaluated in double


def basis(L, z):
    """C_m, C'_m, S_m (m = 1..L) at heights z, clamped to [-1/2, 1/2]: arrays (L, nz).
    get_z_coeff_at (calculate_velocity_operator.cu:452-476); lambda from the oracle."""
    z = np.clip(np.atleast_1d(np.asarray(z, np.float64)), -0.5, 0.5)
    C = np.zeros((L, z.size))
    Cp = np.zeros((L, z.size))
    S = np.zeros((L, z.size))
    for i in range(L):
        k = i + 1
        lam = O.lam(k)
        lz = lam * z
        if k % 2 == 1:
            ch, c = np.cosh(0.5 * lam), np.cos(0.5 * lam)
            C[i] = np.cosh(lz) / ch - np.cos(lz) / c
            Cp[i] = lam * (np.sinh(lz) / ch + np.sin(lz) / c)
        else:
            sh, s = np.sinh(0.5 * lam), np.sin(0.5 * lam)
            C[i] = np.sinh(lz) / sh - np.sin(lz) / s
            Cp[i] = lam * (np.cosh(lz) / sh - np.cos(lz) / s)
        S[i] = np.sin(np.pi * k * (z + 0.5))
    return C, Cp, S


def wavenumbers(M, N, Lx, Ly):
    """kx (mox,), ky (N,) as create_derivative_real_matrix_x/_y (calculate_velocity_operator.cu:
    330-378): row index wraps to negative beyond N/2."""
    mox = M // 2 + 1
    kx = np.arange(mox) * (2 * np.pi / Lx)
    rows = np.arange(N)
    ky = np.where(rows < N // 2 + 1, rows, rows - N) * (2 * np.pi / Ly)
    return kx, ky


def velocity_spectral(state, M, N, Lx, Ly, z):
    """Spectral (u, v, w) at heights z, each (nz, N, mox) complex: the derivative fields of
    calculate_derivatives (calculate_velocity_operator.cu:172-245, add_dc_part :479-494)
    composed with the basis (compose_u_from_derivs :426-450)."""
    th, f, g, F, G = state
    L = f.shape[0]
    kx, ky = wavenumbers(M, N, Lx, Ly)
    KX, KY = np.meshgrid(kx, ky)
    u11 = 1j * KX[None] * f
    u21 = 1j * KY[None] * f
    u31 = (KX ** 2 + KY ** 2)[None] * f
    u12 = 1j * KY[None] * g
    u22 = -1j * KX[None] * g
    u12[:, 0, 0] += F
    u22[:, 0, 0] += G
    C, Cp, S = basis(L, z)
    u = np.einsum("mz,myx->zyx", Cp, u11) + np.einsum("mz,myx->zyx", S, u12)
    v = np.einsum("mz,myx->zyx", Cp, u21) + np.einsum("mz,myx->zyx", S, u22)
    w = np.einsum("mz,myx->zyx", C, u31)
    return u, v, w


def velocity_grid(state, M, N, Lx, Ly, z):
    """(u_x, u_y, u_z) on the M x N grid at heights z, each (nz, N, M): the unnormalised C2R
    of the spectral velocity (cufftExecC2R, calculate_velocity_operator.cu:214-238)."""
    return [np.fft.irfft2(c, s=(N, M)) * (M * N) for c in velocity_spectral(state, M, N, Lx, Ly, z)]


def linspace_levels():
    """The 128 tracer heights: linspace(-PARTICLE_Z_MAX, PARTICLE_Z_MAX) computed in float
    (particle_tracer.cu:152-154, :494-501)."""
    zmin, zmax = np.float32(-ZMAX), np.float32(ZMAX)
    idx = np.arange(ZNUM, dtype=np.float32)
    return (zmin + idx * (zmax - zmin) / np.float32(ZNUM - 1)).astype(np.float64)


def trilinear(grid, Lx, Ly, x, y, z):
    """vel_interpolate (particle_tracer.cu:508-597, PERIODIC build) in FP64 on arrays."""
    ux, uy, uz = grid
    nz, ys, xs = ux.shape
    xg = x * xs / Lx
    yg = y * ys / Ly
    zg = (ZMAX + z) / (2 * ZMAX) * (ZNUM - 1)
    xg = xg - np.floor(xg / xs) * xs
    yg = yg - np.floor(yg / ys) * ys
    ix0 = np.floor(xg).astype(np.int64)
    iy0 = np.floor(yg).astype(np.int64)
    iz0 = np.floor(zg).astype(np.int64)
    ix0 = np.where((ix0 < 0) | (ix0 >= xs), 0, ix0)
    iy0 = np.where((iy0 < 0) | (iy0 >= ys), 0, iy0)
    iz0 = np.clip(iz0, 0, ZNUM - 1)
    ix1 = np.where(ix0 + 1 >= xs, 0, ix0 + 1)
    iy1 = np.where(iy0 + 1 >= ys, 0, iy0 + 1)
    iz1 = np.where(iz0 + 1 >= ZNUM, iz0, iz0 + 1)
    out = []
    for c in (ux, uy, uz):
        wx = xg - ix0
        a00 = (1 - wx) * c[iz0, iy0, ix0] + wx * c[iz0, iy0, ix1]
        a10 = (1 - wx) * c[iz0, iy1, ix0] + wx * c[iz0, iy1, ix1]
        a01 = (1 - wx) * c[iz1, iy0, ix0] + wx * c[iz1, iy0, ix1]
        a11 = (1 - wx) * c[iz1, iy1, ix0] + wx * c[iz1, iy1, ix1]
        wy = yg - iy0
        b0 = (1 - wy) * a00 + wy * a10
        b1 = (1 - wy) * a01 + wy * a11
        wz = zg - iz0
        out.append((1 - wz) * b0 + wz * b1)
    return out


def reflect_lids(z):
    """periodic_boundaries (particle_tracer.cu:722-764, PERIODIC build): x, y unbounded;
    z reflected at +-PARTICLE_Z_MAX, set to the lid beyond twice that; NaN to the top lid."""
    z = np.array(z, np.float64, copy=True)
    top = ~(z <= ZMAX)
    z[top] = np.where(z[top] < 2 * ZMAX, 2 * ZMAX - z[top], ZMAX)
    bot = z < -ZMAX
    z[bot] = np.where(z[bot] > -2 * ZMAX, -2 * ZMAX - z[bot], -ZMAX)
    return z


def rk4_step(grid, Lx, Ly, x, y, z, dt):
    """step_particle_on_GPU with RK4 (particle_tracer.cu:610-687), then the lid reflection."""
    u1 = trilinear(grid, Lx, Ly, x, y, z)
    d1 = [dt * c for c in u1]
    u2 = trilinear(grid, Lx, Ly, x + 0.5 * d1[0], y + 0.5 * d1[1], z + 0.5 * d1[2])
    d2 = [dt * c for c in u2]
    u3 = trilinear(grid, Lx, Ly, x + 0.5 * d2[0], y + 0.5 * d2[1], z + 0.5 * d2[2])
    d3 = [dt * c for c in u3]
    u4 = trilinear(grid, Lx, Ly, x + d3[0], y + d3[1], z + d3[2])
    d4 = [dt * c for c in u4]
    xn = x + (d1[0] + 2 * (d2[0] + d3[0]) + d4[0]) / 6.0
    yn = y + (d1[1] + 2 * (d2[1] + d3[1]) + d4[1]) / 6.0
    zn = z + (d1[2] + 2 * (d2[2] + d3[2]) + d4[2]) / 6.0
    return xn, yn, reflect_lids(zn)


def velocity_points(state, M, N, Lx, Ly, x, y, z):
    """Exact velocity at points: the Fourier sum of the spectral velocity at each point's
    height (z clamped to [-1/2, 1/2]), over the stored half spectrum with weight 1 at kx = 0
    and 2 beyond; real parts only, as a C2R takes them."""
    th, f, g, F, G = state
    L = f.shape[0]
    mox = M // 2 + 1
    x, y, z = (np.asarray(a, np.float64) for a in (x, y, z))
    kx, ky = wavenumbers(M, N, Lx, Ly)
    nzr = np.nonzero(np.any(f != 0, axis=0) | np.any(g != 0, axis=0))
    rows, cols = nzr
    C, Cp, S = basis(L, z)                                   # (L, np)
    KX, KY = kx[cols], ky[rows]
    fm = f[:, rows, cols]                                    # (L, nm)
    gm = g[:, rows, cols]
    A = Cp.T @ fm                                            # (np, nm)
    B = S.T @ gm
    Cc = C.T @ fm
    uk = 1j * KX * A + 1j * KY * B
    vk = 1j * KY * A - 1j * KX * B
    wk = (KX ** 2 + KY ** 2) * Cc
    ph = np.exp(1j * (np.outer(x, KX) + np.outer(y, KY)))
    wgt = np.where(cols == 0, 1.0, 2.0)
    u = (wgt * (ph * uk).real).sum(axis=1) + S.T @ F.real
    v = (wgt * (ph * vk).real).sum(axis=1) + S.T @ G.real
    w = (wgt * (ph * wk).real).sum(axis=1)
    return u, v, w


def rk2_step(state, M, N, Lx, Ly, x, y, z, dt):
    """Midpoint RK2 with the exact velocity, then the lid reflection (the library's spectral
    tracer mode)."""
    u, v, w = velocity_points(state, M, N, Lx, Ly, x, y, z)
    xm, ym, zm = x + 0.5 * dt * u, y + 0.5 * dt * v, z + 0.5 * dt * w
    u, v, w = velocity_points(state, M, N, Lx, Ly, xm, ym, zm)
    return x + dt * u, y + dt * v, reflect_lids(z + dt * w)


def mode_energies(state, M, N, Lx, Ly, z, wz):
    """sum_l wz_l |c_k(z_l)|^2 per stored mode for c = u, v, w: arrays (N, mox)."""
    u, v, w = velocity_spectral(state, M, N, Lx, Ly, z)
    wz = np.asarray(wz, np.float64)[:, None, None]
    return [(wz * np.abs(c) ** 2).sum(axis=0) for c in (u, v, w)]


def energy_spectrum(state, M, N, Lx, Ly, nq):
    """E_h(|k|) in shells of width min(2 pi/Lx, 2 pi/Ly), (1/2)|u_h|^2 integrated over z with
    the nq-point Gauss-Legendre rule, and the totals (E_h, E_z)."""
    zq, wq = O.gauss_legendre(nq)
    eu, ev, ew = mode_energies(state, M, N, Lx, Ly, zq, wq)
    kx, ky = wavenumbers(M, N, Lx, Ly)
    KX, KY = np.meshgrid(kx, ky)
    wk = np.where(np.arange(M // 2 + 1)[None, :] == 0, 1.0, 2.0) * np.ones_like(KX)
    dk = min(2 * np.pi / Lx, 2 * np.pi / Ly)
    shell = np.floor(np.sqrt(KX ** 2 + KY ** 2) / dk + 0.5).astype(int)
    eh = 0.5 * wk * (eu + ev)
    nz = (eu + ev + ew) != 0
    nb = shell[nz].max() + 1 if nz.any() else 1
    E = np.bincount(shell[nz], weights=eh[nz], minlength=nb)
    return E, (eh.sum(), (0.5 * wk * ew).sum())


def energy_spectrum_ref(state, M, N, Lx, Ly, z):
    """reduce_spectrum (energy_spectrum_2d.cu:85-129) of |c_k|^2 for the three velocity
    components at height z: array (3, size), size = max(M/2+1, N/2+1)."""
    mox = M // 2 + 1
    eu, ev, ew = mode_energies(state, M, N, Lx, Ly, [z], [1.0])
    size = max(mox, N // 2 + 1)
    kmx = 2 * np.pi * (mox - 1) / Lx
    kmy = 2 * np.pi * (N // 2) / Ly
    inc = np.sqrt(kmx ** 2 + kmy ** 2) / size
    out = np.zeros((3, size))
    for j in range(N):
        my = 2 * np.pi * (j if j <= N // 2 else N - j) / Ly
        for i in range(mox):
            mx = 2 * np.pi * i / Lx
            a = np.sqrt(mx * mx + my * my)
            for b in range(size):
                if b * inc <= a < (b + 1) * inc:
                    out[0, b] += eu[j, i]
                    out[1, b] += ev[j, i]
                    out[2, b] += ew[j, i]
                    break
    return out


def write_state_file(path, arr, complex_dtype=np.complex64):
    """matrix_folder_writer::write_binary_file (output/matrix_folder_writer.cu:61-103) of one
    field: (L, N, mox) -> dims {mox, N, L}; an (L,) vector -> {1, 1, L}."""
    a = np.asarray(arr)
    dims = [a.shape[2], a.shape[1], a.shape[0]] if a.ndim == 3 else [1, 1, a.shape[0]]
    with open(path, "wb") as fh:
        fh.write(np.array([1, 3] + dims, np.int32).tobytes())
        fh.write(np.ascontiguousarray(a, complex_dtype).tobytes())


def read_state_file(path):
    """read_binary_file (output/matrix_folder_writer.cu:3-58) of a one-matrix file: (dims,
    values as complex128), the element size inferred from the file size."""
    b = open(path, "rb").read()
    nmat, nd = np.frombuffer(b[:8], np.int32)
    assert nmat == 1
    dims = np.frombuffer(b[8:8 + 4 * nd], np.int32).tolist()
    off = 8 + 4 * nd
    n = int(np.prod(dims))
    esz = (len(b) - off) // n
    dt = {8: np.complex64, 16: np.complex128}[esz]
    return dims, np.frombuffer(b[off:], dt).astype(np.complex128)
