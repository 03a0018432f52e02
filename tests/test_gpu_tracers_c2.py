"""Tracers at the 128^2 x 16 configuration and at L = 24, 32, tracer diffusion and the lid
reflection on the GPU (through the C-ABI) against oracle/velocity.py.

- Spectral RK2 at 128^2 x 16 with 1e5 tracers (the point-evaluation instantiation for L <= 16
  at the size the configuration names), FP32 and BF16X3 states; particles move independently,
  so a seeded subset of 2,000 is followed by the FP64 oracle.  L = 24 and 32 run the other two
  instantiations.
- Diffusion: in a zero flow a step adds sigma * N(0, 1) per axis with sigma = sqrt(2 D dt)
  (particle_tracer.cu:196-200, 692-717); after n steps the mean-square displacement per axis
  is 2 D dt n.  Checked on 1e5 tracers (relative standard error of a sample variance:
  sqrt(2 / 1e5) = 0.45 %; bound 3 %), with means within 5 standard errors of 0.
- Lid reflection (periodic_boundaries, particle_tracer.cu:722-764): positions placed beyond
  the lids in a zero flow land where oracle.reflect_lids puts them, in both tracer modes.

Tolerances: positions 2e-5 absolute after a few steps (box lengths ~4)."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import velocity as V
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def _gpu(cfg, precision="f32"):
    from sdc_amd import Stepper
    return Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                   precision=precision)


def _gentle(cfg, seed):
    """A smooth random state scaled to max |u| = 1 on a few levels."""
    M, N, L = cfg["M"], cfg["N"], cfg["L"]
    st = random_state(M, N, L, seed=seed, amp=5e-2)
    kx, ky = V.wavenumbers(M, N, cfg["Lx"], cfg["Ly"])
    KX, KY = np.meshgrid(kx, ky)
    damp = np.exp(-0.3 * np.sqrt(KX ** 2 + KY ** 2))
    st = [a * damp[None] for a in st[:3]] + list(st[3:])
    umax = max(np.abs(c).max() for c in V.velocity_grid(st, M, N, cfg["Lx"], cfg["Ly"], V.linspace_levels()[::16]))
    return [a / umax for a in st]


C2 = dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=2e-4)


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


@pytest.mark.parametrize("cfg,n", [(C2, 100_000),
                                   (dict(M=64, N=64, L=24, Lx=4.0, Ly=3.0, Ra=1e5, Pr=0.7, dt=2e-4), 4000),
                                   (dict(M=64, N=64, L=32, Lx=3.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5), 4000)],
                         ids=["128x128x16-1e5", "64x64x24", "64x64x32"])
@pytest.mark.parametrize("precision", ["f32", "bf16x3"])
def test_spectral_rk2_tracers_vs_oracle(cfg, n, precision):
    """FP32 states run the vertical sums on the FP32 matrix cores, BF16X3 states as BF16 hi/lo
    split products (relative error about 2^-16 per product); both against the FP64 oracle."""
    M, N, L = cfg["M"], cfg["N"], cfg["L"]
    print(f"case: spectral rk2 {n} tracers {M}x{N}x{L} {precision}", flush=True)
    from sdc_amd import Tracer
    state = _gentle(cfg, seed=L + 1)
    steps, dt = 3, 2e-3
    with _gpu(cfg, precision) as G:
        G.set_state(*state)
        T = Tracer(G, n, buffer_size=8, mode="spectral_rk2", seed=6)
        T.init_positions("random")
        x0, y0, z0 = T.positions()
        for _ in range(steps):
            T.step_particles(dt)
        x1, y1, z1 = T.positions()
        T.close()
    assert np.all(np.isfinite(x1)) and np.all(np.isfinite(z1))
    sel = np.random.default_rng(0).choice(n, size=min(n, 2000), replace=False)
    x, y, z = (a[sel].astype(np.float64) for a in (x0, y0, z0))
    for _ in range(steps):
        x, y, z = V.rk2_step(state, M, N, cfg["Lx"], cfg["Ly"], x, y, z, dt)
    err = max(np.abs(x1[sel] - x).max(), np.abs(y1[sel] - y).max(), np.abs(z1[sel] - z).max())
    print("max position error", err)
    assert err < 2e-5, err


@pytest.mark.parametrize("mode", ["rk4_trilinear", "spectral_rk2"])
def test_tracer_diffusion_mean_square_displacement(mode):
    print("case: tracer diffusion", mode, flush=True)
    from sdc_amd import Tracer
    cfg = dict(M=32, N=32, L=8, Lx=4.0, Ly=4.0, Ra=5000.0, Pr=1.0, dt=1e-3)
    n, D, dt, steps = 100_000, 0.1, 1e-3, 4
    with _gpu(cfg) as G:  # zero state: no flow
        T = Tracer(G, n, buffer_size=8, mode=mode, diff_coeff=D, seed=11)
        T.init_positions("center")
        x0, y0, z0 = T.positions()
        for _ in range(steps):
            T.step_particles(dt)
        x1, y1, z1 = T.positions()
        T.close()
    msd_want = 2 * D * dt * steps
    se = np.sqrt(msd_want / n)
    for a0, a1 in ((x0, x1), (y0, y1), (z0, z1)):
        d = a1.astype(np.float64) - a0.astype(np.float64)
        assert abs(d.mean()) < 5 * se, (d.mean(), se)
        msd = (d ** 2).mean()
        assert abs(msd / msd_want - 1) < 0.03, (msd, msd_want)
    dx, dy = x1.astype(np.float64) - x0, y1.astype(np.float64) - y0
    assert abs(np.corrcoef(dx, dy)[0, 1]) < 0.02


@pytest.mark.parametrize("mode", ["rk4_trilinear", "spectral_rk2"])
def test_tracer_lid_reflection_vs_oracle(mode):
    print("case: tracer lid reflection", mode, flush=True)
    from sdc_amd import Tracer
    cfg = dict(M=32, N=32, L=8, Lx=4.0, Ly=4.0, Ra=5000.0, Pr=1.0, dt=1e-3)
    z = np.array([0.0, 0.3, 0.4989, 0.499, 0.4995, 0.5, 0.7, 0.9, 0.997, 1.5, 3.0,
                  -0.4989, -0.4995, -0.5, -0.7, -0.997, -1.2, -3.0, 0.25], np.float32)
    n = len(z)
    x = np.linspace(0.1, 3.9, n).astype(np.float32)
    y = np.linspace(3.7, 0.2, n).astype(np.float32)
    with _gpu(cfg) as G:  # zero state: the step moves nothing, only the lids act
        T = Tracer(G, n, buffer_size=4, mode=mode, seed=1)
        T.set_positions(x, y, z)
        T.step_particles(1e-3)
        x1, y1, z1 = T.positions()
        T.close()
    want = V.reflect_lids(z.astype(np.float64))
    assert np.array_equal(x1, x) and np.array_equal(y1, y)
    assert np.all(np.isfinite(z1))
    assert np.abs(z1 - want).max() < 1e-6, np.c_[z, z1, want]
