"""GPU parity of libsdc_rbc (through the C-ABI) against the FP64 CPU oracle.

Tolerances: FP64 path <= 1e-10 relative after 100 steps (and ~1e-12 per operator call);
FP32 path relative errors as stated per test."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sdc_amd.initial import kept_mask, low_mode_theta, random_theta
from tests.helpers import random_state, rel_err

pytestmark = pytest.mark.gpu

C1 = dict(M=32, N=32, L=8, Lx=4.032, Ly=4.032, Ra=5000.0, Pr=1.0, dt=1e-3)


def _gpu(cfg, precision="f64", **kw):
    from sdc_amd import Stepper
    return Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                   precision=precision, **kw)


def _orc(cfg, nthreads=8):
    return O.Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                     nthreads=nthreads)


def _cmp(a, b, keep):
    errs = []
    for i in range(3):
        errs.append(rel_err(a[i][:, keep], b[i][:, keep]))
    for i in (3, 4):
        errs.append(rel_err(a[i], b[i]) if np.abs(b[i]).max() > 0 else float(np.abs(a[i]).max()))
    return errs


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def test_c1_operators_fp64():
    print("case: C1 fp64 operators", flush=True)
    st = random_state(C1["M"], C1["N"], C1["L"], seed=11, amp=1e-2)
    keep = kept_mask(C1["M"], C1["N"])
    with _gpu(C1) as G:
        o = _orc(C1)
        e = _cmp(G.calculate_nonlinear(*st), o.nonlinear(*st), keep)
        assert max(e) < 1e-12, e
        e = _cmp(G.calculate_inverse(*st), o.inverse_B(*st), keep)
        assert max(e) < 1e-14, e
        e = _cmp(G.calculate_implicit(*st), o.implicit(*st), keep)
        assert max(e) < 1e-13, e


def test_c1_100_steps_fp64_parity():
    """North-star criterion: FP64 fields within 1e-10 relative after 100 steps (C1)."""
    print("case: C1 fp64 100 steps", flush=True)
    st = low_mode_theta(C1["M"], C1["N"], C1["L"], seed=1, amp=1e-3)
    # a developed-looking start: add small f, g, F, G so every path is exercised
    extra = random_state(C1["M"], C1["N"], C1["L"], seed=2, amp=1e-3)
    st = [st[0] + extra[0], extra[1], extra[2], extra[3], extra[4]]
    keep = kept_mask(C1["M"], C1["N"])
    with _gpu(C1) as G:
        o = _orc(C1)
        G.set_state(*st)
        o.set_state(*st)
        G.step_time(100)
        o.step(100)
        e = _cmp(G.get_state(), o.get_state(), keep)
        print("C1 fp64 100-step rel errors", e)
        assert max(e) < 1e-10, e
        nu_g, nu_o = G.nusselt(), o.nusselt()
        assert np.abs(nu_g - nu_o).max() < 1e-10


def test_c1_fp32_vs_fp64_oracle():
    print("case: C1 fp32 50 steps", flush=True)
    st = low_mode_theta(C1["M"], C1["N"], C1["L"], seed=1, amp=1e-1)
    keep = kept_mask(C1["M"], C1["N"])
    with _gpu(C1, "f32") as G:
        o = _orc(C1)
        G.set_state(*st)
        o.set_state(*st)
        G.step_time(50)
        o.step(50)
        e = _cmp(G.get_state(), o.get_state(), keep)
        print("C1 fp32 50-step rel errors", e)
        assert max(e[:3]) < 1e-4, e


@pytest.mark.parametrize("M,N,L", [(64, 32, 4), (32, 64, 6), (128, 128, 16)])
def test_nonlinear_shapes_fp64(M, N, L):
    print(f"case: nonlinear fp64 {M}x{N}x{L}", flush=True)
    cfg = dict(M=M, N=N, L=L, Lx=3.3, Ly=4.1, Ra=2e4, Pr=0.7, dt=2e-4)
    st = random_state(M, N, L, seed=M + N + L, amp=1e-2, decay=0.05)
    keep = kept_mask(M, N)
    with _gpu(cfg) as G:
        o = _orc(cfg)
        e = _cmp(G.calculate_nonlinear(*st), o.nonlinear(*st), keep)
        assert max(e) < 1e-11, e


@pytest.mark.parametrize("M,N,L", [(64, 32, 4), (64, 128, 8), (256, 64, 8), (128, 128, 16), (256, 256, 8)])
def test_group_fft_kernels_fp32(M, N, L):
    """FP32 nonlinear term through the natural-layout group transforms (x-rows and y-FFTs of 64,
    128 and 256 points; N = 32 keeps the generic y-FFT) against the generic LDS kernels and the
    FP64 oracle; 2e-5 relative."""
    print(f"case: group-FFT fp32 {M}x{N}x{L}", flush=True)
    cfg = dict(M=M, N=N, L=L, Lx=3.3, Ly=4.1, Ra=2e4, Pr=0.7, dt=2e-4)
    st = random_state(M, N, L, seed=M + 2 * N + L, amp=1e-2, decay=0.05)
    keep = kept_mask(M, N)
    ref = _orc(cfg).nonlinear(*st)
    with _gpu(cfg, "f32") as G:
        group = G.calculate_nonlinear(*st)
        os.environ["RBC_GENERIC_FFT"] = "1"
        try:
            generic = G.calculate_nonlinear(*st)
        finally:
            del os.environ["RBC_GENERIC_FFT"]
    e1, e2, e3 = _cmp(group, ref, keep), _cmp(generic, ref, keep), _cmp(group, generic, keep)
    print("group vs oracle", e1, "generic vs oracle", e2, "group vs generic", e3)
    assert max(e1) < 2e-5 and max(e2) < 2e-5 and max(e3) < 2e-5


def test_bundled_512_state_fp32_and_fp64():
    """The reference's bundled 512^2 x 2 state (FP32) through the nonlinear operator."""
    print("case: bundled 512x512x2 state", flush=True)
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "data_state_kept.npz"))
    M, N, L = 512, 512, 2
    keep = kept_mask(M, N)
    st = []
    for n in ("theta", "f", "g"):
        a = np.zeros((L, N, M // 2 + 1), np.complex128)
        a[:, keep] = d[n]
        st.append(a)
    st += [d["F"].astype(np.complex128), d["G"].astype(np.complex128)]
    cfg = dict(M=M, N=N, L=L, Lx=2 * 2.016 * 2, Ly=2 * 2.016 * 2, Ra=5000.0, Pr=1.0, dt=1e-3)
    o = _orc(cfg)
    ref = o.nonlinear(*st)
    with _gpu(cfg, "f64") as G:
        e = _cmp(G.calculate_nonlinear(*st), ref, keep)
        assert max(e) < 1e-11, e
    with _gpu(cfg, "f32") as G:
        fast = G.calculate_nonlinear(*st)
        os.environ["RBC_GENERIC_FFT"] = "1"
        try:
            generic = G.calculate_nonlinear(*st)
        finally:
            del os.environ["RBC_GENERIC_FFT"]
    e = _cmp(fast, ref, keep)
    print("512^2x2 fp32 (512-point kernels) nonlinear rel errors vs oracle", e)
    eg = _cmp(generic, ref, keep)
    print("512^2x2 fp32 (generic LDS kernels) nonlinear rel errors vs oracle", eg)
    # FP32 on a developed state: the f/g terms carry k^2..k^3 derivative weights
    assert max(e) < 1e-4 and max(eg) < 1e-4, (e, eg)
    ef = _cmp(fast, generic, keep)
    print("512-point vs generic kernels", ef)
    assert max(ef) < 1e-4, ef


def test_c2_fp64_five_steps():
    print("case: C2 fp64 5 steps", flush=True)
    cfg = dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=2e-4)
    st = random_theta(128, 128, 16, seed=7, amp=1e-3)
    keep = kept_mask(128, 128)
    with _gpu(cfg) as G:
        o = _orc(cfg, nthreads=16)
        G.set_state(*st)
        o.set_state(*st)
        G.step_time(5)
        o.step(5)
        e = _cmp(G.get_state(), o.get_state(), keep)
        assert max(e[:3]) < 1e-11, e


def test_gpu_state_invariants_after_steps():
    print("case: invariants fp32", flush=True)
    st = low_mode_theta(C1["M"], C1["N"], C1["L"], seed=3, amp=1e-1)
    keep = kept_mask(C1["M"], C1["N"])
    with _gpu(C1, "f32") as G:
        G.set_state(*st)
        G.step_time(10)
        th, f, g, F, G_ = G.get_state()
        N = C1["N"]
        for a in (th, f, g):
            assert np.all(a[:, ~keep] == 0)
            assert np.all(a[:, 0, 0].imag == 0)
            for r in range(1, N // 2):
                assert np.array_equal(a[:, r, 0], np.conj(a[:, N - r, 0]))
        assert np.all(F.imag == 0) and np.all(G_.imag == 0)


def test_512_fast_path_vs_oracle_L4():
    """The 512-point FP32 kernels (x rows and y columns) against the FP64 oracle."""
    print("case: 512x512x4 fp32 nonlinear vs oracle", flush=True)
    cfg = dict(M=512, N=512, L=4, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5)
    st = random_state(512, 512, 4, seed=21, amp=1e-2, decay=0.02)
    keep = kept_mask(512, 512)
    ref = _orc(cfg, nthreads=16).nonlinear(*st)
    with _gpu(cfg, "f32") as G:
        e = _cmp(G.calculate_nonlinear(*st), ref, keep)
        print("512^2x4 fp32 rel errors", e)
        assert max(e) < 2e-5, e


@pytest.mark.parametrize("M,N", [(512, 256), (256, 512)])
def test_512_one_side_fp32_vs_oracle(M, N):
    """Grids with 512 points on one side only: the 512-point kernels need both sides (their
    hybrid rows are tile-major), so these run the group / generic kernels on the row-major
    layout with the 512-point row stride."""
    print(f"case: {M}x{N}x4 fp32 nonlinear vs oracle", flush=True)
    cfg = dict(M=M, N=N, L=4, Lx=4.0, Ly=4.0 * N / M, Ra=1e6, Pr=0.7, dt=5e-5)
    st = random_state(M, N, 4, seed=23, amp=1e-2, decay=0.02)
    keep = kept_mask(M, N)
    ref = _orc(cfg, nthreads=16).nonlinear(*st)
    with _gpu(cfg, "f32") as G:
        e = _cmp(G.calculate_nonlinear(*st), ref, keep)
        print(f"{M}x{N}x4 fp32 rel errors", e)
        assert max(e) < 2e-5, e


def test_c3_fp32_vs_fp64_gpu_steps():
    """Full 512^2 x 32 workload: FP32 fast path vs the FP64 GPU path (itself pinned to the
    oracle above) after 20 steps."""
    print("case: C3 fp32 vs fp64, 20 steps", flush=True)
    cfg = dict(M=512, N=512, L=32, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5)
    st = random_theta(512, 512, 32, seed=3, amp=1e-3)
    keep = kept_mask(512, 512)
    with _gpu(cfg, "f64") as G:
        G.set_state(*st)
        G.step_time(20)
        ref = G.get_state()
    with _gpu(cfg, "f32") as G:
        G.set_state(*st)
        G.step_time(20)
        out = G.get_state()
    e = _cmp(out, ref, keep)
    print("C3 fp32 vs fp64 20-step rel errors", e)
    assert max(e[:3]) < 1e-4, e


def test_fp32_matrix_core_vs_valu_and_oracle():
    """FP32 vertical transforms on the matrix cores against the VALU kernels and the oracle."""
    print("case: fp32 mfma vs valu, 64x64x16", flush=True)
    cfg = dict(M=64, N=64, L=16, Lx=3.0, Ly=3.5, Ra=3e4, Pr=0.7, dt=2e-4)
    st = random_state(64, 64, 16, seed=31, amp=1e-2, decay=0.05)
    keep = kept_mask(64, 64)
    ref = _orc(cfg).nonlinear(*st)
    with _gpu(cfg, "f32") as G:
        mma = G.calculate_nonlinear(*st)
    os.environ["RBC_NO_MFMA"] = "1"
    try:
        with _gpu(cfg, "f32") as G:
            valu = G.calculate_nonlinear(*st)
    finally:
        del os.environ["RBC_NO_MFMA"]
    e1, e2, e3 = _cmp(mma, ref, keep), _cmp(valu, ref, keep), _cmp(mma, valu, keep)
    print("fp32 mfma vs oracle", e1, "valu vs oracle", e2, "mfma vs valu", e3)
    assert max(e1) < 2e-5 and max(e2) < 2e-5 and max(e3) < 2e-5


def test_bf16x3_vertical_transforms():
    """BF16 hi/lo split matrix-core vertical transforms (FP32 accumulation): nonlinear term vs
    the oracle and 20 steps of the 512^2 x 32 workload vs the FP64 GPU path."""
    print("case: bf16x3 nonlinear 64x64x16 and C3 20 steps", flush=True)
    cfg = dict(M=64, N=64, L=16, Lx=3.0, Ly=3.5, Ra=3e4, Pr=0.7, dt=2e-4)
    st = random_state(64, 64, 16, seed=31, amp=1e-2, decay=0.05)
    keep = kept_mask(64, 64)
    ref = _orc(cfg).nonlinear(*st)
    with _gpu(cfg, "bf16x3") as G:
        e = _cmp(G.calculate_nonlinear(*st), ref, keep)
    print("bf16x3 nonlinear rel errors", e)
    assert max(e) < 2e-4, e
    cfg3 = dict(M=512, N=512, L=32, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5)
    st3 = random_theta(512, 512, 32, seed=3, amp=1e-3)
    keep3 = kept_mask(512, 512)
    with _gpu(cfg3, "f64") as G:
        G.set_state(*st3)
        G.step_time(20)
        r3 = G.get_state()
    with _gpu(cfg3, "bf16x3") as G:
        G.set_state(*st3)
        G.step_time(20)
        o3 = G.get_state()
    e3 = _cmp(o3, r3, keep3)
    print("C3 bf16x3 vs fp64 20-step rel errors", e3)
    assert max(e3[:3]) < 1e-4, e3
