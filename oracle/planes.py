"""FP64 numpy restatement of the reference's plane operators (gpu_flow/code/operator/
freeze_operator.cu, masktemp.cu, masking.cu): every vertical mode plane of theta (and f, g for
the freeze) goes to physical space, is combined point by point under an integer mask, scaled
by 1/(M N) and transformed back.

TEST INFRASTRUCTURE: only tests/ import this module.

Fields are complex arrays (L, N, M/2+1) in the reference layout (x fastest); masks are integer
arrays (N, M) indexed [row, col] (1 = the operator leaves the point alone).
"""
import numpy as np


def plane_mask(shape, M, N, param, cxf=0.5, cyf=0.5, fp32=False):
    """init_mask_physical_space_rectangle / _circle of freeze_operator.cu (centre (0.5 M, 0.5 N)),
    masktemp.cu (circle centre (0.2 M, 0.3 N)) and masking.cu (centre (c_y M, c_x N)), in the
    build's real type: rectangle -> 1 within `param` of a side (as a fraction of the side);
    circle -> 1 outside radius param * min(M, N) / 2."""
    rt = np.float32 if fp32 else np.float64
    col = np.arange(M)[None, :].astype(rt)
    row = np.arange(N)[:, None].astype(rt)
    if shape == "rectangle":
        w = rt(param)
        m = (col < w * rt(M)) | (col > (rt(1) - w) * rt(M)) | (row < w * rt(N)) | (row > (rt(1) - w) * rt(N))
    else:
        dx = col - rt(cxf) * rt(M)
        dy = row - rt(cyf) * rt(N)
        r = rt(param) * rt(0.5) * rt(min(M, N))
        m = (dx * dx + dy * dy) > r * r
    return m.astype(np.int32)


def c2r(a, M, N):
    """Unnormalised inverse transform of one (N, M/2+1) plane (the reference's C2R)."""
    return np.fft.irfft2(a, s=(N, M)) * (M * N)


def r2c(x):
    return np.fft.rfft2(x)


def capture(fields, M, N):
    """freeze_mask: the physical planes of the fields, (nf, L, N, M)."""
    return np.array([[c2r(f[m], M, N) for m in range(f.shape[0])] for f in fields])


def apply(fields, op, mask, M, N, frozen=None, factor=1.0):
    """apply_mask of freeze_operator.cu ("freeze": mask ? u : frozen), masktemp.cu ("heat":
    mask ? u : u + 1.5/100) or julie_do_the_thing of masking.cu ("scale": mask ? u : u * factor),
    each times 1/(M N), then R2C."""
    out = []
    keep = mask == 1
    for i, f in enumerate(fields):
        g = np.empty_like(f)
        for m in range(f.shape[0]):
            u = c2r(f[m], M, N)
            if op == "freeze":
                v = np.where(keep, u, frozen[i, m])
            elif op == "heat":
                v = np.where(keep, u, u + 1.5 / 100)
            else:
                v = np.where(keep, u, u * factor)
            g[m] = r2c(v / (M * N))
        out.append(g)
    return out
