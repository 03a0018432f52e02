"""Parity at the full 512 x 512 x 32 workload through a size-independent property.

A horizontal Fourier mode of amplitude 1e-12 evolves linearly: its products with itself and
with the other seeded modes are smaller by many orders of magnitude (also with the velocity
that Ra = 1e6 drives from theta) and land on other wavenumbers, so its trajectory depends on
(kx, ky), the vertical basis, Ra, Pr and dt, and not on the grid size.
Three single modes are seeded at 512^2 x 32 and at 32^2 x 32 (the same physical wavenumbers:
both grids keep them) and stepped 40 times; the 512^2 run, in every precision, must carry
the same theta, f and g values on those modes as the FP64 32^2 run (whose path the oracle
tests pin), and nothing near the modes' size anywhere else.  Tolerances: FP64 1e-11, FP32 and BF16X3
1e-4 (the linear part runs in FP32 in both; the vertical transforms only touch the products).
"""
import numpy as np
import pytest

from sdc_amd.initial import gaussian_complex, kept_mask
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

PHYS = (4.0, 4.0, 1e6, 0.7, 5e-5)  # Lx, Ly, Ra, Pr, dt of the 512^2 x 32 workload
L = 32
MODES = [(1, 1), (-2, 3), (3, 0)]  # (ky, kx)
AMP = 1e-12
STEPS = 40


def _state(M, N):
    mox = M // 2 + 1
    th = np.zeros((L, N, mox), np.complex128)
    z = gaussian_complex(17, L * len(MODES)).reshape(len(MODES), L) * AMP
    for i, (ky, kx) in enumerate(MODES):
        th[:, ky % N, kx] = z[i]
        if kx == 0:  # the kx = 0 column holds both ky and -ky
            th[:, (-ky) % N, 0] = np.conj(z[i])
    assert all(kept_mask(M, N)[ky % N, kx] for ky, kx in MODES)
    zeros = np.zeros((L, N, mox), np.complex128)
    return [th, zeros.copy(), zeros.copy(), np.zeros(L, np.complex128), np.zeros(L, np.complex128)]


def _run(M, N, precision):
    from sdc_amd import Stepper
    with Stepper(M, N, L, *PHYS, precision=precision) as S:
        S.set_state(*_state(M, N))
        S.step_time(STEPS)
        return S.get_state()


def _on_modes(fields, N):
    return np.concatenate([fields[c][:, ky % N, kx] for c in range(3) for ky, kx in MODES])


@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("f32", 1e-4), ("bf16x3", 1e-4)])
def test_single_modes_evolve_as_on_a_small_grid(precision, tol):
    print(f"case: single modes 512x512x32 {precision} vs 32x32x32 f64, {STEPS} steps", flush=True)
    ref = _run(32, 32, "f64")
    out = _run(512, 512, precision)
    a, b = _on_modes(out, 512), _on_modes(ref, 32)
    assert np.abs(b).max() > 0.1 * AMP  # the modes are alive, f and g driven from theta
    e = rel_err(a, b)
    print("rel err on the seeded modes", e)
    assert e < tol, e
    # everything else stays at the order of the products (amplitude^2), far below the modes
    rest = [f.copy() for f in out[:3]]
    for r in rest:
        for ky, kx in MODES:
            r[:, ky % 512, kx] = 0
            if kx == 0:
                r[:, (-ky) % 512, 0] = 0
    leak = max(np.abs(r).max() for r in rest)
    print("largest other coefficient", leak)
    assert leak < 1e-4 * np.abs(a).max(), leak
    assert np.abs(out[3]).max() < 1e-4 * AMP and np.abs(out[4]).max() < 1e-4 * AMP
