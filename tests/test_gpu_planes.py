"""Plane operators (operator/freeze_operator.cu, masktemp.cu, masking.cu) on the GPU against
oracle/planes.py, and the driver's -freeze_circle / -temp_circle loop logic (main.cu:569-590).
Tolerances: FP64 1e-12, FP32 2e-6 relative to the field's largest kept coefficient."""
import os
import subprocess

import numpy as np
import pytest

from oracle import planes as PL
from sdc_amd.initial import kept_mask
from tests.helpers import random_state, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = (64, 32, 8, 4.0, 2.0, 2e4, 0.7, 1e-3)


def _kept(a, keep):
    b = a.copy()
    b[:, ~keep] = 0
    return b


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_freeze_heat_and_theta_mask_vs_oracle(precision, tol):
    print("case: plane operators", precision, flush=True)
    from sdc_amd import Stepper
    M, N = CFG[0], CFG[1]
    keep = kept_mask(M, N)
    a = random_state(M, N, 8, seed=61, amp=1e-2)
    b = random_state(M, N, 8, seed=62, amp=1e-2)
    fp32 = precision != "f64"
    with Stepper(*CFG, precision=precision) as S:
        S.set_state(*a)
        S.apply_freeze()                    # before any freeze: nothing happens
        assert all(np.array_equal(x, y) for x, y in zip(S.get_state(), S.get_state()))
        S.freeze("circle", 0.8)             # captures a's planes
        S.set_state(*b)
        S.apply_freeze()
        got_freeze = S.get_state()
        S.set_state(*b)
        base = S.get_state()                # b as stored in the build's precision
        S.temp_mask("rectangle", 0.1)
        S.apply_temp_mask()
        got_heat = S.get_state()
        S.set_state(*b)
        S.mask_theta(0.4, 0.6, 0.3, 1.1)
        got_scale = S.get_state()
    ka = [_kept(x, keep) for x in a[:3]]
    kb = [_kept(x, keep) for x in b[:3]]
    frozen = PL.capture(ka, M, N)
    want = PL.apply(kb, "freeze", PL.plane_mask("circle", M, N, 0.8, fp32=fp32), M, N, frozen=frozen)
    e = [rel_err(g[:, keep], w[:, keep]) for g, w in zip(got_freeze[:3], want)]
    print("freeze", e)
    assert max(e) < tol, e
    want = PL.apply(kb[:1], "heat", PL.plane_mask("rectangle", M, N, 0.1, fp32=fp32), M, N)[0]
    e = rel_err(got_heat[0][:, keep], want[:, keep])
    print("heat", e)
    assert e < tol and all(np.array_equal(x, y) for x, y in zip(got_heat[1:], base[1:]))
    # masking.cu's circle: centre (c_y M, c_x N) = (0.6 M, 0.4 N)
    want = PL.apply(kb[:1], "scale", PL.plane_mask("circle", M, N, 0.3, 0.6, 0.4, fp32=fp32), M, N, factor=1.1)[0]
    e = rel_err(got_scale[0][:, keep], want[:, keep])
    print("scale", e)
    assert e < tol


def test_driver_freeze_and_temp_circle(tmp_path):
    """-freeze_circle n and -temp_circle n: from step n on, the operators run before each step
    (the freeze captures at step n after its own no-op application, masktemp prepares its
    width-0.1 rectangle likewise), equal to the same loop through the library API."""
    print("case: driver plane operators", flush=True)
    from sdc_amd import Stepper
    d = str(tmp_path / "run")
    run = os.path.join(ROOT, "sdc_amd", "rbc_run")
    args = ["-nmodes", "32", "32", "8", "-clength", "4.032", "4.032", "-ra", "5000", "-pr", "1", "-dt", "0.001",
            "-t", "0.02", "-nfiles", "1", "-init_test", "-precision", "f64", "-freeze_circle", "5",
            "-temp_circle", "8", "-dir", d]
    r = subprocess.run([run] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    with Stepper(32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3, precision="f64") as S:
        S.init_state("test")
        for i in range(1, 21):
            S.apply_freeze()
            if i == 5:
                S.freeze("circle", 0.8)
            S.apply_temp_mask()
            if i == 8:
                S.temp_mask("rectangle", 0.1)
            S.step_time(1)
        names = [str(tmp_path / f"api_{n}.bin") for n in ("theta", "f", "g", "F", "G")]
        S.write_state_files(names)
    for n, ref in zip(("theta", "f", "g", "F", "G"), names):
        assert open(os.path.join(d, f"{n}_00000020.bin"), "rb").read() == open(ref, "rb").read(), n
