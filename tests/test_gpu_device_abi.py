"""Device-resident C-ABI calls (rbc_*_device): state exchange and the three operators on
device buffers in the reference layout, element type per precision.  They run the same
kernels as the host-buffer calls, so their results must equal those bit for bit, and the
host calls are pinned to the oracle; the operators are also compared with the oracle here.
Device memory comes from hipMalloc (tests/hipmem.py); the library works on its own stream, so
the test synchronises the device after uploads and the context before reading."""
import numpy as np
import pytest

from oracle import oracle as O
from sdc_amd.initial import kept_mask
from tests.helpers import random_state, rel_err

pytestmark = pytest.mark.gpu

CASES = [(dict(M=32, N=32, L=8, Lx=4.032, Ly=4.032, Ra=5000.0, Pr=1.0, dt=1e-3), "f64", 1e-12),
         (dict(M=64, N=32, L=16, Lx=4.0, Ly=2.5, Ra=3e4, Pr=0.7, dt=2e-4), "f32", 2e-5),
         # BF16 hi/lo split vertical transforms (relative error about 2^-16 per product)
         (dict(M=64, N=32, L=16, Lx=4.0, Ly=2.5, Ra=3e4, Pr=0.7, dt=2e-4), "bf16x3", 2e-4)]


def _dev(arrays, precision):
    from tests.hipmem import DeviceArray
    dt = np.complex128 if precision == "f64" else np.complex64
    return [DeviceArray(a.astype(dt)) for a in arrays]


def _host(bufs):
    return [b.download().astype(np.complex128) for b in bufs]


def _like(st, precision, fill=1.0):
    from tests.hipmem import DeviceArray
    dt = np.complex128 if precision == "f64" else np.complex64
    return [DeviceArray.full(a.shape, dt, fill) for a in st]


@pytest.mark.parametrize("cfg,precision,tol", CASES, ids=["32x32x8-f64", "64x32x16-f32", "64x32x16-bf16x3"])
def test_device_state_and_operators(cfg, precision, tol):
    from sdc_amd import Stepper
    from tests.hipmem import synchronize
    M, N, L = cfg["M"], cfg["N"], cfg["L"]
    print(f"case: device ABI {M}x{N}x{L} {precision}", flush=True)
    st = random_state(M, N, L, seed=M + L, amp=1e-2, decay=0.05)
    if precision != "f64":  # the values a complex64 buffer can hold
        st = [a.astype(np.complex64).astype(np.complex128) for a in st]
    keep = kept_mask(M, N)
    o = O.Stepper(M, N, L, cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"], nthreads=8)
    ref = [o.nonlinear(*st), o.inverse_B(*st), o.implicit(*st)]
    o.close()
    with Stepper(M, N, L, cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"], precision=precision) as S:
        host = [S.calculate_nonlinear(*st), S.calculate_inverse(*st), S.calculate_implicit(*st)]
        din = _dev(st, precision)
        synchronize()
        calls = [S.calculate_nonlinear_device, S.calculate_inverse_device, S.calculate_implicit_device]
        for k, call in enumerate(calls):
            out = _like(st, precision)  # the removed band must come back as zeros
            synchronize()
            call([t.ptr for t in din], [t.ptr for t in out])
            S.synchronize()
            got = _host(out)
            for a, b in zip(got, host[k]):
                assert np.array_equal(a, b), f"operator {k}: device call differs from host call"
            for a in got[:3]:
                assert np.all(a[:, ~keep] == 0)
            errs = [rel_err(got[i][:, keep], ref[k][i][:, keep]) for i in range(3)]
            assert max(errs) < (tol if k == 0 else 10 * tol), (k, errs)
        # state round trip and steps on device buffers equal the host-buffer path
        S.set_state_device(*[t.ptr for t in din])
        S.step_time(4)
        out = _like(st, precision, fill=0.0)
        synchronize()
        S.get_state_device(*[t.ptr for t in out])
        S.synchronize()
        dev_steps = _host(out)
        S.set_state(*st)
        S.reset_history()
        S.step_time(4)
        host_steps = S.get_state()
    for a, b in zip(dev_steps, host_steps):
        assert np.array_equal(a, b)


def test_velocity_at_device_equals_host_call():
    """rbc_velocity_at_device on device arrays equals rbc_velocity_at on host arrays bit for bit
    (same plan and kernels), across calls that grow and then reuse the context's scratch."""
    from sdc_amd import Stepper
    from tests.hipmem import DeviceArray, synchronize
    print("case: velocity_at_device", flush=True)
    cfg = (64, 64, 16, 4.0, 3.0, 2e4, 0.7, 1e-3)
    st = random_state(64, 64, 16, seed=77, amp=5e-2, decay=0.1)
    rng = np.random.default_rng(8)
    with Stepper(*cfg, precision="f32") as S:
        S.set_state(*st)
        for n in (3000, 50_000, 700):
            x = rng.uniform(0, 4.0, n).astype(np.float32)
            y = rng.uniform(0, 3.0, n).astype(np.float32)
            z = rng.uniform(-0.5, 0.5, n).astype(np.float32)
            host = S.velocity_at(x, y, z)
            dx, dy, dz = (DeviceArray(a) for a in (x, y, z))
            out = [DeviceArray(np.zeros(n, np.float32)) for _ in range(3)]
            synchronize()
            S.velocity_at_device(n, dx.ptr, dy.ptr, dz.ptr, *[o.ptr for o in out])
            S.synchronize()
            for a, b in zip(out, host):
                assert np.array_equal(a.download(), b), n
