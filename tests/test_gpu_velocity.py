"""GPU checks of the state files, the velocity operator, both tracer modes and the energy
spectra (through the C-ABI) against the FP64 restatement in oracle/velocity.py.

Tolerances: FP64 spectra 1e-12 relative; FP32 velocity 2e-6 of the field's maximum; FP32
tracer positions 2e-5 absolute after a few steps (box lengths ~4)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import velocity as V
from sdc_amd.initial import kept_mask
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

CFG = dict(M=64, N=64, L=8, Lx=4.0, Ly=3.0, Ra=2e4, Pr=0.7, dt=1e-3)


def _gpu(cfg, precision="f32"):
    from sdc_amd import Stepper
    return Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                   precision=precision)


@pytest.fixture(scope="module")
def state():
    O.build()
    return random_state(CFG["M"], CFG["N"], CFG["L"], seed=5, amp=5e-2)


def _args():
    return CFG["M"], CFG["N"], CFG["Lx"], CFG["Ly"]


@pytest.fixture(scope="module")
def gentle(state):
    """The test state scaled to max |u| = 1 and smoothed, so that a tracer moves a fraction of
    a grid cell per step and float rounding is not amplified by the flow's strain."""
    kx, ky = V.wavenumbers(*_args())
    KX, KY = np.meshgrid(kx, ky)
    damp = np.exp(-0.3 * np.sqrt(KX ** 2 + KY ** 2))
    st = [a * damp[None] for a in state[:3]] + list(state[3:])
    umax = max(np.abs(c).max() for c in V.velocity_grid(st, *_args(), V.linspace_levels()[::8]))
    return [a / umax for a in st]


def test_state_files_bundled_round_trip(tmp_path):
    """The reference's bundled 512^2 x 2 FP32 state written in its format, read by the library
    (dims check, make_real) and written back byte-identically."""
    print("case: state files 512x512x2", flush=True)
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "data_state_kept.npz"))
    M, N, L = 512, 512, 2
    keep = kept_mask(M, N)
    fields = []
    for n in ("theta", "f", "g"):
        a = np.zeros((L, N, M // 2 + 1), np.complex64)
        a[:, keep] = d[n]
        fields.append(a)
    fields += [d["F"], d["G"]]
    names = ["theta", "f", "g", "F", "G"]
    src = [str(tmp_path / f"{n}_in.bin") for n in names]
    dst = [str(tmp_path / f"{n}_out.bin") for n in names]
    for p, a in zip(src, fields):
        V.write_state_file(p, a, np.complex64)
    from sdc_amd import Stepper
    with Stepper(M, N, L, 8.064, 8.064, 5000.0, 1.0, 1e-3, precision="f32") as G:
        G.read_state_files(src)
        got = G.get_state()
        for a, b in zip(got, fields):
            assert np.array_equal(a, b.astype(np.complex128))
        G.write_state_files(dst)
    for p, q in zip(src, dst):
        assert open(p, "rb").read() == open(q, "rb").read()
    # a wrong size is refused with the reference's message
    with Stepper(256, 256, L, 8.064, 8.064, 5000.0, 1.0, 1e-3, precision="f32") as G2:
        with pytest.raises(Exception, match="right dimension"):
            G2.read_state_files(src)


def test_history_files_restart_continues_ab2(tmp_path, state):
    """Restart from state + AB2 history files reproduces the uninterrupted run exactly; a
    restart from the state alone takes an Euler step and differs."""
    print("case: restart with history", flush=True)
    names = ["theta", "f", "g", "F", "G"]
    sp = [str(tmp_path / f"{n}.bin") for n in names]
    hp = [str(tmp_path / f"{n}_hist.bin") for n in names]
    with _gpu(CFG, "f64") as A:
        A.set_state(*state)
        A.step_time(3)
        A.write_state_files(sp)
        A.write_state_files(hp, which=1)
        A.step_time(4)
        ref = A.get_state()
    with _gpu(CFG, "f64") as B:
        B.read_state_files(sp)
        B.read_state_files(hp, which=1)
        B.step_time(4)
        got = B.get_state()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    with _gpu(CFG, "f64") as C:
        C.read_state_files(sp)
        C.step_time(4)
        euler = C.get_state()
    assert np.abs(euler[0] - ref[0]).max() > 0


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_velocity_grid_vs_oracle(state, precision):
    print("case: velocity grid", precision, flush=True)
    z = V.linspace_levels()[::9]
    with _gpu(CFG, precision) as G:
        G.set_state(*state)
        got = G.velocity_grid(z)
    ref = V.velocity_grid(state, *_args(), z)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 2e-6 * np.abs(b).max(), np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_velocity_at_points_vs_oracle(state, precision):
    print("case: velocity at points", precision, flush=True)
    rng = np.random.default_rng(3)
    n = 1000
    x = rng.uniform(-CFG["Lx"], 2 * CFG["Lx"], n).astype(np.float32)
    y = rng.uniform(0, CFG["Ly"], n).astype(np.float32)
    z = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    with _gpu(CFG, precision) as G:
        G.set_state(*state)
        got = G.velocity_at(x, y, z)
        # at grid nodes the exact sum equals the C2R grid
        zl = np.array([-0.25, 0.3])
        grid = G.velocity_grid(zl)
        ix, iy, iz = np.arange(0, 64, 7), np.arange(0, 64, 5)[:10], np.array([0, 1] * 5)[:10]
        ix = ix[:10]
        nodes = G.velocity_at((ix * CFG["Lx"] / CFG["M"]).astype(np.float32),
                              (iy * CFG["Ly"] / CFG["N"]).astype(np.float32), zl[iz].astype(np.float32))
    ref = V.velocity_points(state, *_args(), x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 2e-5 * np.abs(b).max(), np.abs(a - b).max() / np.abs(b).max()
    for a, g in zip(nodes, grid):
        assert np.abs(a - g[iz, iy, ix]).max() < 2e-5 * np.abs(g).max()


def test_rk4_tracer_vs_oracle(gentle):
    state = gentle
    print("case: rk4 trilinear tracer", flush=True)
    from sdc_amd import Tracer
    n, steps, dt = 2000, 5, 2e-3
    with _gpu(CFG, "f32") as G:
        G.set_state(*state)
        T = Tracer(G, n, buffer_size=16, mode="rk4_trilinear", seed=2)
        T.init_positions("random")
        x0, y0, z0 = T.positions()
        for _ in range(steps):
            T.step_particles(dt)
        xs, ys, zs = T.get_particle_positions(clear=False)
        x1, y1, z1 = T.positions()
        T.close()
    # random init in the box (particle_tracer.cu:441-451)
    assert 0 < x0.min() and x0.max() <= CFG["Lx"] and 0 < y0.min() and y0.max() <= CFG["Ly"]
    assert -0.5 < z0.min() and z0.max() <= 0.5
    assert xs.shape == (steps, n) and np.array_equal(xs[0], x0)
    grid = V.velocity_grid(state, *_args(), V.linspace_levels())
    x, y, z = x0.astype(np.float64), y0.astype(np.float64), z0.astype(np.float64)
    for _ in range(steps):
        x, y, z = V.rk4_step(grid, CFG["Lx"], CFG["Ly"], x, y, z, dt)
    err = max(np.abs(x1 - x).max(), np.abs(y1 - y).max(), np.abs(z1 - z).max())
    assert err < 2e-5, err


def test_spectral_rk2_tracer_vs_oracle(gentle):
    state = gentle
    print("case: spectral rk2 tracer", flush=True)
    from sdc_amd import Tracer
    n, steps, dt = 1500, 4, 2e-3
    with _gpu(CFG, "f32") as G:
        G.set_state(*state)
        T = Tracer(G, n, buffer_size=8, mode="spectral_rk2", seed=4)
        T.init_positions("random")
        x0, y0, z0 = T.positions()
        for _ in range(steps):
            T.step_particles(dt)
        x1, y1, z1 = T.positions()
        T.close()
    x, y, z = x0.astype(np.float64), y0.astype(np.float64), z0.astype(np.float64)
    for _ in range(steps):
        x, y, z = V.rk2_step(state, *_args(), x, y, z, dt)
    err = max(np.abs(x1 - x).max(), np.abs(y1 - y).max(), np.abs(z1 - z).max())
    assert err < 2e-5, err


def test_tracer_ring_buffer_and_moments(state):
    """is_buffer_full / get_particle_positions(Clear) semantics and the FP64 moments."""
    print("case: tracer ring", flush=True)
    from sdc_amd import Tracer, RbcError
    n = 64
    with _gpu(CFG, "f32") as G:
        G.set_state(*state)
        T = Tracer(G, n, buffer_size=4, mode="rk4_trilinear")
        x, y, z = T.positions()  # the constructor's centre placement
        assert np.all(x == np.float32(0.5 * CFG["Lx"])) and np.all(z == 0)
        T.init_positions("pairs")
        px, py, _ = T.positions()
        assert np.array_equal(px[0::2], px[1::2]) and np.array_equal(py[0::2], py[1::2])
        for _ in range(3):
            T.step_particles(1e-3)
        with pytest.raises(RbcError, match="filled"):
            T.step_particles(1e-3)
        cur = T.positions()
        xs, _, _ = T.get_particle_positions(clear=True)
        assert xs.shape == (3, n)
        assert T.info() == (0, 3)
        assert np.array_equal(T.positions()[0], cur[0])
        T.step_particles(1e-3)
        m = T.moments()
        x, y, _ = T.positions()
        xd, yd = x.astype(np.float64), y.astype(np.float64)
        assert abs(m[0] - xd.mean()) < 1e-12 and abs(m[1] - yd.mean()) < 1e-12
        assert abs(m[2] - ((xd ** 2 + yd ** 2).mean() - xd.mean() ** 2 - yd.mean() ** 2)) < 1e-9
        dx = np.abs(x[1::2] - x[0::2]); dy = np.abs(y[1::2] - y[0::2])
        dx = np.where(dx > 0.5 * CFG["Lx"], CFG["Lx"] - dx, dx)
        dy = np.where(dy > 0.5 * CFG["Ly"], CFG["Ly"] - dy, dy)
        assert abs(m[3] - np.mean(dx.astype(np.float64) ** 2 + dy.astype(np.float64) ** 2)) < 1e-9
        T.close()


def test_energy_spectra_vs_oracle(state):
    print("case: energy spectra", flush=True)
    with _gpu(CFG, "f64") as G:
        G.set_state(*state)
        E, tot = G.energy_spectrum()
        E3 = G.energy_spectrum_ref(0.2)
        nq = G.nq
    Er, totr = V.energy_spectrum(state, *_args(), nq)
    E3r = V.energy_spectrum_ref(state, *_args(), 0.2)
    assert E.shape == Er.shape
    assert np.abs(E - Er).max() < 1e-12 * Er.max()
    assert abs(tot[0] - totr[0]) < 1e-12 * totr[0] and abs(tot[1] - totr[1]) < 1e-12 * totr[1]
    assert E3.shape == E3r.shape
    assert np.abs(E3 - E3r).max() < 1e-12 * E3r.max()
