"""The simulation driver (rbc_run, ≙ main.cu) and the initial states on the GPU.

- rbc_init_state("test") equals the oracle's restatement of init_test (util/init.cu:263-301)
  after make_real, and 100 FP64 steps from it match the oracle (1e-10), so a run can start from
  an input the reference itself would produce without random numbers.
- rbc_init_state("random"): every kept theta coefficient within +-amplitude, mean ~0, variance
  amplitude^2 / 3 (init_random's uniform distribution), f, g, F, G zero, Hermitian column.
- rbc_run with a reference-style argument line (stale README names -delta_t / -final_t,
  -init_test, -nfiles) writes theta_%08d.bin .. G_%08d.bin at iCeil(t / dt) / nfiles step
  multiples, byte-identical to the same run through the library API; -mean_F applies
  force_mean before every step; tracer tracks (.trkss) and second moments are written.
The driver runs as a child process (the program itself, not a shell)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from sdc_amd.initial import kept_mask
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "sdc_amd", "rbc_run")
C1 = (32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3)


def _run(args, timeout=120):
    print("rbc_run", " ".join(args), flush=True)
    r = subprocess.run([RUN] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def test_init_test_state_vs_oracle_and_100_steps():
    print("case: init_test state", flush=True)
    from sdc_amd import Stepper
    keep = kept_mask(32, 32)
    ref = O.init_test_state(32, 32, 8)
    for a in ref[:3]:
        O.dealias(a)
        O.make_real(a)
    with Stepper(*C1, precision="f64") as S:
        S.init_state("test")
        got = S.get_state()
        for a, b in zip(got, ref):
            assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max() + 1e-300
        S.step_time(100)
        out = S.get_state()
    o = O.Stepper(*C1, nthreads=8)
    o.set_state(*ref)
    o.step(100)
    r = o.get_state()
    o.close()
    e = [rel_err(out[i][:, keep], r[i][:, keep]) for i in range(3)] + [rel_err(out[3], r[3])]
    print("init_test 100-step errors", e)
    assert max(e) < 1e-10, e


def test_init_random_distribution():
    print("case: init random", flush=True)
    from sdc_amd import Stepper
    amp = 1e-6
    with Stepper(64, 64, 8, 4.0, 4.0, 2e4, 0.7, 1e-3, precision="f64") as S:
        S.init_state("random", seed=5, amplitude=amp)
        th, f, g, F, G = S.get_state()
        S.init_state("random", seed=5, amplitude=amp)
        again = S.get_state()[0]
        S.init_state("random", seed=6, amplitude=amp)
        other = S.get_state()[0]
    keep = kept_mask(64, 64)
    v = th[:, keep]
    assert np.array_equal(th, again) and not np.array_equal(th, other)
    assert np.all(np.abs(v.real) <= amp) and np.all(np.abs(v.imag) <= amp)
    body = th[:, 1:, 1:][:, keep[1:, 1:]]  # away from the kx = 0 column that make_real rewrites
    n = body.size
    for part in (body.real, body.imag):
        assert abs(part.mean()) < 5 * amp / np.sqrt(3 * n)
        assert abs(part.var() / (amp ** 2 / 3) - 1) < 0.05
    assert not np.any(f) and not np.any(g) and not np.any(F) and not np.any(G)
    assert np.all(th[:, ~keep] == 0) and np.all(th[:, 0, 0].imag == 0)
    for r in range(1, 32):
        assert np.array_equal(th[:, r, 0], np.conj(th[:, 64 - r, 0]))


def _api_files(tmp, precision, steps, every, mean_F=0.0, init="test"):
    from sdc_amd import Stepper
    # an FP32 build reads its real arguments as float (CUDA_FLOAT_REAL), so the API run uses
    # the float-rounded values
    cfg = C1 if precision == "f64" else tuple(v if isinstance(v, int) else float(np.float32(v)) for v in C1)
    out = {}
    with Stepper(*cfg, precision=precision) as S:
        S.init_state(init)
        for i in range(1, steps + 1):
            if mean_F:
                S.force_mean(mean_F)
            S.step_time(1)
            if i % every == 0:
                names = [os.path.join(tmp, f"api_{n}_{i:08d}.bin") for n in ("theta", "f", "g", "F", "G")]
                S.write_state_files(names)
                out[i] = names
    return out


@pytest.mark.parametrize("precision,mean_F", [("f64", 0.0), ("f32", 0.05), ("bf16x3", 0.0)])
def test_driver_files_equal_library_run(tmp_path, precision, mean_F):
    print("case: driver files", precision, mean_F, flush=True)
    d = str(tmp_path / "run")
    args = ["-nmodes", "32", "32", "8", "-clength", "4.032", "4.032", "-ra", "5000", "-pr", "1",
            "-delta_t", "0.001", "-final_t", "0.0305", "-nfiles", "3", "-init_test", "-precision", precision,
            "-dir", d]
    if mean_F:
        args += ["-mean_F", str(mean_F)]
    log = _run(args)
    assert "number of iterations: 31" in log
    assert "number of iterations: 31" in open(os.path.join(d, "simulation.log")).read()
    api = _api_files(str(tmp_path), precision, 30, 10, mean_F=mean_F)
    assert sorted(api) == [10, 20, 30]
    for i, names in api.items():
        for n, ref in zip(("theta", "f", "g", "F", "G"), names):
            got = open(os.path.join(d, f"{n}_{i:08d}.bin"), "rb").read()
            assert got == open(ref, "rb").read(), (n, i)
    if mean_F:  # F_0 holds the forced value before each step, so it differs from the unforced run
        F30 = np.frombuffer(open(os.path.join(d, "F_00000030.bin"), "rb").read()[20:], np.complex64)
        assert abs(F30[0].real) > 0


def test_driver_tracks_and_second_moment(tmp_path):
    print("case: driver tracks", flush=True)
    d = str(tmp_path / "trk")
    args = ["-nmodes", "32", "32", "8", "-clength", "4.032", "4.032", "-ra", "5000", "-pr", "1", "-dt", "0.001",
            "-t", "0.3", "-nfiles", "1", "-init_test", "-precision", "f32", "-npart", "64", "-le", "0",
            "-part_center_sigma", "-write_tracks_every", "2", "-calc_secondmoment", "-dir", d]
    _run(args)
    raw = open(os.path.join(d, "particles_le0e+00_00000300.trkss"), "rb").read()
    assert raw[:6] == b"TRKSS" + b"a"
    ntr, fs, dl = struct.unpack("<3I", raw[6:18])
    desc = raw[18:18 + dl].decode()
    t0, t1, nfr = struct.unpack("<3I", raw[18 + dl:30 + dl])
    assert ntr == 64 and fs == 1 and "Le=0e+00" in desc
    # the ring of 256 slots is read out once (255 steps), every 2nd frame kept
    assert nfr == 128 and t0 == 0 and t1 == 254
    body = raw[30 + dl:]
    assert len(body) == nfr * (4 + 3 * 4 * ntr)
    lines = [ln for ln in open(os.path.join(d, "secondmoment_0.dat")) if not ln.startswith("#")]
    assert len(lines) == 300
    t, m2 = map(float, lines[-1].split("\t"))
    assert abs(t - 0.3) < 1e-6 and m2 > 0


def test_characteristic_velocity_vs_oracle_and_extract_tool(tmp_path):
    """<|u|> over 10 heights (main_extract_vel.cu:182-218) against the FP64 velocity of
    oracle/velocity.py, and the rbc_extract_vel tool's data-point file."""
    print("case: characteristic velocity", flush=True)
    from oracle import velocity as V
    from sdc_amd import Stepper
    from tests.helpers import random_state
    cfg = (64, 32, 8, 4.0, 2.5, 2e4, 0.7, 1e-3)
    st = random_state(64, 32, 8, seed=12, amp=5e-2, decay=0.05)
    z = (np.float32(-0.5) + np.arange(10, dtype=np.float32) * np.float32(1.0) / np.float32(9)).astype(np.float64)
    u, v, w = V.velocity_grid(st, 64, 32, 4.0, 2.5, z)
    want = np.sqrt(u ** 2 + v ** 2 + w ** 2).mean()
    with Stepper(*cfg, precision="f32") as S:
        S.set_state(*st)
        got = S.characteristic_velocity(10)
        names = [str(tmp_path / f"{n}.bin") for n in ("theta", "f", "g", "F", "G")]
        S.write_state_files(names)
    print("U", got, "oracle", want)
    assert abs(got / want - 1) < 1e-5
    out = str(tmp_path / "charvel.dat")
    tool = os.path.join(ROOT, "sdc_amd", "rbc_extract_vel")
    r = subprocess.run([tool, "-clength", "4", "2.5", "-files"] + names +
                       ["-write_charvel_nz", "1", "-outfile_charvel", out, "-description", "test"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = open(out).read().splitlines()
    assert lines[:3] == ["# Data point writer", "# test", "# Data ordering: t, val"]
    t, U = lines[3].split("\t")
    assert float(t) == 0.0 and abs(float(U) / want - 1) < 1e-5
