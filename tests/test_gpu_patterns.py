"""Pattern starts (rbc_init_pattern; main.cu:326-352, util/init.cu) on the GPU through the C-ABI:
vertical mode 0 of theta against oracle/patterns.py (the unnormalised R2C of the restated plane),
the other modes against the plain random start of the same seed, f, g, F, G zero; and the
driver's -init_<pattern> -init_mode options against the same start through the library.
Tolerances: FP64 1e-12, FP32 and BF16X3 (FP32 storage; the start involves no vertical
transform) 2e-6, relative to the largest kept coefficient."""
import os
import subprocess

import numpy as np
import pytest

from oracle import patterns as PT
from sdc_amd.initial import kept_mask
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = (64, 32, 8, 4.0, 2.0, 2e4, 0.7, 1e-3)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 2e-6), ("bf16x3", 2e-6)])
@pytest.mark.parametrize("kind", list(PT.KINDS))
def test_pattern_start_vs_oracle(kind, precision, tol):
    print("case: pattern start", kind, precision, flush=True)
    from sdc_amd import Stepper
    M, N, L = CFG[0], CFG[1], CFG[2]
    keep = kept_mask(M, N)
    seed = 9
    with Stepper(*CFG, precision=precision) as S, Stepper(*CFG, precision=precision) as R:
        S.init_pattern(kind, seed=seed)  # -init_mode default Lx / 2
        th, f, g, F, G = S.get_state()
        R.init_state("random", seed=seed, amplitude=1e-6)
        th_r = R.get_state()[0]
    mode = PT.rectangle_mode(CFG[3], M) if kind == "rectangle" else int(CFG[3] / 2.0)
    want = PT.spectrum(kind, M, N, mode, 1e-6, seed)
    e = rel_err(th[0][keep], want[keep])
    print("mode 0", e)
    assert e < tol, e
    assert np.all(th[:, ~keep] == 0)
    e = rel_err(th[1:][:, keep], th_r[1:][:, keep])
    print("modes >= 1", e)
    assert e < tol, e
    assert not np.any(f) and not np.any(g) and not np.any(F) and not np.any(G)


def test_driver_pattern_options_equal_library(tmp_path):
    """-init_giantspiral -init_mode 3 (-init_cr given too: the later option of main.cu:90-96
    wins) writes the same state files as rbc_init_pattern + the same steps through the library."""
    print("case: driver pattern start", flush=True)
    from sdc_amd import Stepper
    d = str(tmp_path / "run")
    args = ["-nmodes", "32", "32", "8", "-clength", "4.032", "4.032", "-ra", "5000", "-pr", "1", "-dt", "0.001",
            "-t", "0.0045", "-nfiles", "1", "-init_cr", "-init_giantspiral", "-init_mode", "3", "-seed", "4",
            "-precision", "f64", "-dir", d]
    r = subprocess.run([os.path.join(ROOT, "sdc_amd", "rbc_run")] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    with Stepper(32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3, precision="f64") as S:
        S.init_pattern("giantspiral", init_mode=3, seed=4)
        S.step_time(5)
        names = [str(tmp_path / f"api_{n}.bin") for n in ("theta", "f", "g", "F", "G")]
        S.write_state_files(names)
    for n, ref in zip(("theta", "f", "g", "F", "G"), names):
        assert open(os.path.join(d, f"{n}_00000005.bin"), "rb").read() == open(ref, "rb").read(), n
