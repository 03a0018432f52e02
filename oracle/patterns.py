"""Oracle (test infrastructure only): the pattern starts of the reference's main.cu:326-352,
restated from gpu_flow/code/util/init.cu in numpy, FP64.

Each pattern is a real plane of amplitude amp on the M x N grid (point (i, j) at x = i / M,
y = j / N, fractions of the box); the start is theta = the random start (all levels), with
vertical mode 0 replaced by the unnormalised 2-D R2C of the plane (init.cu:312-318: cuFFT R2C,
here numpy.fft.rfft2, the same (N, M/2+1) layout and no scaling).  The reference draws the
rectangle's noise columns and the isr phase from C rand(); the library draws them from
SplitMix64 on (seed, point), restated here (splitmix64_mix), so the comparison is exact.

Functions:
  splitmix64_mix(x)          the SplitMix64 output function of state x (+ the golden gamma)
  uniform(seed, idx)         the library's uniform in [0, 1) of (seed, idx)
  value(kind, x, y, mode, amp)  the deterministic patterns at any point (x, y)
  plane(kind, M, N, mode, amp, seed)  the pattern on the grid, [j][i]
  spectrum(...)              rfft2 of the plane, [ky row][kx]
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
KINDS = {"rectangle": 1, "dislocation": 2, "isr": 3, "sv": 4, "cr": 5, "giantspiral": 6}


def splitmix64_mix(x):
    """SplitMix64 (Steele, Lea, Flood 2014): z = x + gamma, then the two xor-shift-multiply
    rounds and the final xor-shift.  x: uint64 array."""
    with np.errstate(over="ignore"):
        z = np.asarray(x, np.uint64) + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, idx):
    """(mix(seed * 0x100000001B3 + idx) >> 11) * 2^-53, the library's uniform draw."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * np.uint64(0x100000001B3) + np.asarray(idx, np.uint64)
    return (splitmix64_mix(x) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def value(kind, x, y, mode, amp, M=None, N=None, phase=0.0):
    """The pattern at points (x, y) (fractions of the box).  M, N give the grid-dependent core
    positions of dislocation and cr; phase is isr's random phase."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    tp = 2.0 * np.pi
    if k in (1, 3):  # init.cu:70-129: cos(2 pi (i x_mode / M + j y_mode / N)), y_mode = 0
        return amp * np.cos(tp * mode * x + (phase if k == 3 else 0.0))
    if k == 2:  # init.cu:22-68
        dx, dy = 1.0 / M, 1.0 / N
        x1 = (M / 2.0) * dx + dx / 2.0
        y1 = (N / 4.0) * dy + dy / 2.0
        y2 = (3.0 * N / 4.0) * dy + dy / 2.0
        return amp * np.cos(mode * tp * x - np.arctan2(y - y1, x - x1) + np.arctan2(y - y2, x - x1))
    if k == 4:  # init.cu:131-165, x_frac = 0.2
        xf = 0.2
        xl = 0.5 * (1.0 - xf)
        xr = 1.0 - xl
        ph = -0.5 * np.sin(tp * y) - x * x
        ph = ph + np.where(x > xr, 1.0, np.where(x > xl, 0.5 * (1.0 - np.cos(np.pi * (x - xl) / xf)), 0.0))
        return amp * np.cos(tp * (mode * y + ph))
    if k == 5:  # init.cu:167-213, x_frac = 0.8
        dx, dy = 1.0 / M, 1.0 / N
        xf = 0.8
        x1 = 0.5 * ((1.0 - xf) * M + 1) * dx
        x2 = 1.0 - x1
        y1 = 0.5 * (N + 1) * dy
        return amp * np.cos(mode * tp * y - np.arctan2(y - y1, x - x1) + np.arctan2(y - y1, x - x2))
    if k == 6:  # init.cu:215-242
        r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
        return amp * np.cos(tp * mode * r - np.arctan2(y - 0.5, x - 0.5) + np.arctan2(y - 0.5, x - 0.8))
    raise ValueError(kind)


def rectangle_mode(Lx, M):
    """main.cu:330: init_rectangle(theta, (Lx / (M - 1)) (M - 5) / 2, 0), truncated to int."""
    return int((Lx / (M - 1.0)) * (M - 5.0) / 2.0)


def plane(kind, M, N, mode, amp=1e-6, seed=0):
    """The pattern on the grid, shape (N, M) = [j][i]."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    i = np.arange(M)[None, :]
    j = np.arange(N)[:, None]
    x, y = i / M, j / N
    if k == 3:
        ph = 2.0 * np.pi * float(uniform(seed, np.uint64(0xFFFFFFFFFFFFFFFF)))
        return np.broadcast_to(value(3, x, y, mode, amp, phase=ph), (N, M)).copy()
    v = np.broadcast_to(value(k, x, y, mode, amp, M, N), (N, M)).copy()
    if k == 1:  # the columns i < M / 8 hold uniform noise in [-amp, amp]
        e = (np.arange(N)[:, None] * M + np.arange(M)[None, :]).astype(np.uint64)
        noise = (2.0 * uniform(seed, 2 * e) - 1.0) * amp
        v[:, : M // 8] = noise[:, : M // 8]
    return v


def spectrum(kind, M, N, mode, amp=1e-6, seed=0):
    """Unnormalised R2C of the plane (init.cu:312-318), [ky row][kx], shape (N, M/2+1)."""
    return np.fft.rfft2(plane(kind, M, N, mode, amp, seed))
