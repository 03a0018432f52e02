"""Penalization (PENALIZED builds, nonlinear_operator_rayleigh_noslip.cu:1097-1343, masks
:1652-1740) on the GPU against the oracle's restatement of the reference's pairwise summands.

Tolerances: FP64 nonlinear term 1e-12 and 5 FP64 steps 1e-10; FP32 nonlinear term 2e-5; all
relative to the field's largest kept coefficient.  eta = 0.05 makes the mask terms dominate."""
import numpy as np
import pytest

from oracle import oracle as O
from sdc_amd.initial import kept_mask
from tests.helpers import random_state, rel_err

pytestmark = pytest.mark.gpu

SHAPES = ["rectangle", "circle", "diskarray", "everywhere"]


def _cmp(a, b, keep):
    errs = [rel_err(a[i][:, keep], b[i][:, keep]) for i in range(3)]
    for i in (3, 4):
        errs.append(rel_err(a[i], b[i]) if np.abs(b[i]).max() > 0 else float(np.abs(a[i]).max()))
    return errs


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


@pytest.mark.parametrize("shape", SHAPES)
def test_penalized_nonlinear_term_vs_oracle(shape):
    from sdc_amd import Stepper
    M, N, L, Lx, Ly, Pr, eta = 64, 32, 8, 4.0, 2.7, 0.7, 0.05
    print(f"case: penalization {shape} {M}x{N}x{L}", flush=True)
    st = random_state(M, N, L, seed=17, amp=1e-2, decay=0.05)
    keep = kept_mask(M, N)
    o = O.Stepper(M, N, L, Lx, Ly, 2e4, Pr, 2e-4, nthreads=8)
    o.set_penalization(eta, shape)
    ref = o.nonlinear(*st)
    o.close()
    for precision, tol in (("f64", 1e-12), ("f32", 2e-5)):
        with Stepper(M, N, L, Lx, Ly, 2e4, Pr, 2e-4, precision=precision) as S:
            S.set_penalization(eta, shape)
            e = _cmp(S.calculate_nonlinear(*st), ref, keep)
        print(precision, shape, e)
        assert max(e) < tol, (precision, e)


def test_penalized_steps_and_custom_mask():
    """5 FP64 steps with the rectangle mask vs the oracle; the same mask given as an array
    (RBC_MASK_CUSTOM) gives the same state bit for bit; eta = 0 turns the terms off."""
    from sdc_amd import Stepper
    M, N, L, Lx, Ly, Pr, eta = 32, 32, 8, 4.032, 4.032, 1.0, 0.05
    print("case: penalized steps", flush=True)
    st = random_state(M, N, L, seed=23, amp=1e-2)
    keep = kept_mask(M, N)
    o = O.Stepper(M, N, L, Lx, Ly, 5000.0, Pr, 1e-3, nthreads=8)
    o.set_penalization(eta, "rectangle")
    o.set_state(*st)
    o.step(5)
    ref = o.get_state()
    o.close()
    rect = np.zeros((N, M))
    rect[:3, :] = rect[-3:, :] = 1.0
    rect[:, :3] = rect[:, -3:] = 1.0
    outs = []
    for how in ("shape", "array"):
        with Stepper(M, N, L, Lx, Ly, 5000.0, Pr, 1e-3, precision="f64") as S:
            if how == "shape":
                S.set_penalization(eta, "rectangle")
            else:
                S.set_penalization(eta, "custom", mask=rect)
            S.set_state(*st)
            S.step_time(5)
            outs.append(S.get_state())
    e = _cmp(outs[0], ref, keep)
    print("penalized 5 steps", e)
    assert max(e) < 1e-10, e
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    with Stepper(M, N, L, Lx, Ly, 5000.0, Pr, 1e-3, precision="f64") as S:
        plain = S.calculate_nonlinear(*st)
        S.set_penalization(eta, "rectangle")
        S.set_penalization(0.0, "rectangle")
        again = S.calculate_nonlinear(*st)
    for a, b in zip(plain, again):
        assert np.array_equal(a, b)


def test_penalized_l32_fp32_vs_oracle():
    """L = 32 (the work arrays reuse the level-space buffers) in FP32 and BF16X3 vs the oracle."""
    from sdc_amd import Stepper
    M, N, L, Lx, Ly, Pr, eta = 64, 64, 32, 4.0, 4.0, 0.7, 0.1
    print("case: penalization L=32", flush=True)
    st = random_state(M, N, L, seed=29, amp=1e-2, decay=0.05)
    keep = kept_mask(M, N)
    o = O.Stepper(M, N, L, Lx, Ly, 1e6, Pr, 5e-5, nthreads=16)
    o.set_penalization(eta, "circle")
    ref = o.nonlinear(*st)
    o.close()
    for precision, tol in (("f32", 2e-5), ("bf16x3", 2e-4)):
        with Stepper(M, N, L, Lx, Ly, 1e6, Pr, 5e-5, precision=precision) as S:
            S.set_penalization(eta, "circle")
            e = _cmp(S.calculate_nonlinear(*st), ref, keep)
        print(precision, e)
        assert max(e) < tol, e


def test_driver_penalization_option_equals_library(tmp_path):
    """rbc_run -penalization circle -eta 0.05 (the -DCIRCLE PENALIZED build) writes the same state
    files as rbc_set_penalization(0.05, circle) and the same steps through the library."""
    import os
    import subprocess
    from sdc_amd import Stepper
    print("case: driver penalization", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    d = str(tmp_path / "run")
    args = ["-nmodes", "32", "32", "8", "-clength", "4.032", "4.032", "-ra", "5000", "-pr", "1", "-dt", "0.001",
            "-t", "0.0045", "-nfiles", "1", "-init_test", "-eta", "0.05", "-penalization", "circle",
            "-precision", "f64", "-dir", d]
    r = subprocess.run([os.path.join(root, "sdc_amd", "rbc_run")] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    with Stepper(32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3, precision="f64") as S:
        S.init_state("test")
        S.set_penalization(0.05, "circle")
        S.step_time(5)
        names = [str(tmp_path / f"api_{n}.bin") for n in ("theta", "f", "g", "F", "G")]
        S.write_state_files(names)
        plain = Stepper(32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3, precision="f64")
        plain.init_state("test")
        plain.step_time(5)
        th_plain = plain.get_state()[0]
        plain.close()
        th_pen = S.get_state()[0]
    for n, ref in zip(("theta", "f", "g", "F", "G"), names):
        assert open(os.path.join(d, f"{n}_00000005.bin"), "rb").read() == open(ref, "rb").read(), n
    assert np.abs(th_pen - th_plain).max() > 0  # the penalization acted
