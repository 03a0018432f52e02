"""GPU parity at the headline's vertical resolution (L = 32) and in boxes with Lx != Ly,
through the C-ABI, against the FP64 CPU oracle.

L = 32 pins what only that size exercises: the 104 (FP64) and 96 (FP32) level rules, the
32-term class bound of the group update, three K-slabs of the matrix-core vertical
transforms and the 32 x 32 parity-split factor records.  Lx != Ly pins the |k|^2 grouping
key of the factor records (kx^2 and ky^2 enter with different weights).

Tolerances: FP64 operators 1e-12 (nonlinear), 1e-14 (inverse B), 1e-13 (implicit), FP64
steps 1e-10; FP32 nonlinear 2e-5, BF16X3 nonlinear 2e-4, FP32 linear operators 2e-6 (factors
rounded to FP32); all relative to the field's largest coefficient."""
import numpy as np
import pytest

from oracle import oracle as O
from sdc_amd.initial import kept_mask, random_theta
from tests.helpers import random_state, rel_err

pytestmark = pytest.mark.gpu


def _gpu(cfg, precision="f64"):
    from sdc_amd import Stepper
    return Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                   precision=precision)


def _orc(cfg):
    return O.Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                     nthreads=16)


def _cmp(a, b, keep):
    errs = [rel_err(a[i][:, keep], b[i][:, keep]) for i in range(3)]
    for i in (3, 4):
        errs.append(rel_err(a[i], b[i]) if np.abs(b[i]).max() > 0 else float(np.abs(a[i]).max()))
    return errs


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


L32 = [dict(M=64, N=64, L=32, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5),
       dict(M=128, N=128, L=32, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5)]


@pytest.mark.parametrize("cfg", L32, ids=["64x64x32", "128x128x32"])
def test_l32_operators_all_precisions(cfg):
    M, N, L = cfg["M"], cfg["N"], cfg["L"]
    print(f"case: L=32 operators {M}x{N}x{L}", flush=True)
    st = random_state(M, N, L, seed=M + 32, amp=1e-2, decay=0.05)
    keep = kept_mask(M, N)
    o = _orc(cfg)
    ref_n, ref_b, ref_i = o.nonlinear(*st), o.inverse_B(*st), o.implicit(*st)
    o.close()
    with _gpu(cfg, "f64") as G:
        print("fp64 levels", G.nq)
        en = _cmp(G.calculate_nonlinear(*st), ref_n, keep)
        eb = _cmp(G.calculate_inverse(*st), ref_b, keep)
        ei = _cmp(G.calculate_implicit(*st), ref_i, keep)
    print("fp64 nonlinear", en, "inverse B", eb, "implicit", ei)
    assert max(en) < 1e-12 and max(eb) < 1e-14 and max(ei) < 1e-13, (en, eb, ei)
    with _gpu(cfg, "f32") as G:
        print("fp32 levels", G.nq)
        en = _cmp(G.calculate_nonlinear(*st), ref_n, keep)
        eb = _cmp(G.calculate_inverse(*st), ref_b, keep)
        ei = _cmp(G.calculate_implicit(*st), ref_i, keep)
    print("fp32 nonlinear", en, "inverse B", eb, "implicit", ei)
    assert max(en) < 2e-5 and max(eb) < 2e-6 and max(ei) < 2e-6, (en, eb, ei)
    with _gpu(cfg, "bf16x3") as G:
        en = _cmp(G.calculate_nonlinear(*st), ref_n, keep)
    print("bf16x3 nonlinear", en)
    assert max(en) < 2e-4, en


@pytest.mark.parametrize("cfg,steps", [(L32[0], 5), (L32[1], 3)], ids=["64x64x32", "128x128x32"])
def test_l32_fp64_steps_vs_oracle(cfg, steps):
    """Euler start then AB2 steps at L = 32, FP64, from a seeded theta plus small f, g, F, G."""
    M, N, L = cfg["M"], cfg["N"], cfg["L"]
    print(f"case: L=32 fp64 {steps} steps {M}x{N}x{L}", flush=True)
    th = random_theta(M, N, L, seed=9, amp=1e-2)
    extra = random_state(M, N, L, seed=10, amp=1e-3, decay=0.05)
    st = [th[0] + extra[0], extra[1], extra[2], extra[3], extra[4]]
    keep = kept_mask(M, N)
    o = _orc(cfg)
    o.set_state(*st)
    o.step(steps)
    ref = o.get_state()
    nu_o = o.nusselt()
    o.close()
    with _gpu(cfg, "f64") as G:
        G.set_state(*st)
        G.step_time(steps)
        out = G.get_state()
        nu_g = G.nusselt()
    e = _cmp(out, ref, keep)
    print(f"L=32 fp64 {steps}-step rel errors", e)
    assert max(e) < 1e-10, e
    assert np.abs(nu_g - nu_o).max() < 1e-10 * max(1.0, np.abs(nu_o).max())


NONSQ = [dict(M=64, N=32, L=16, Lx=4.0, Ly=2.7, Ra=3e4, Pr=0.7, dt=2e-4),
         dict(M=32, N=64, L=8, Lx=2.2, Ly=5.1, Ra=2e4, Pr=1.3, dt=5e-4)]


@pytest.mark.parametrize("cfg", NONSQ, ids=["64x32x16", "32x64x8"])
def test_non_square_box_operators_and_steps_fp64(cfg):
    M, N, L = cfg["M"], cfg["N"], cfg["L"]
    print(f"case: non-square box {M}x{N}x{L} Lx={cfg['Lx']} Ly={cfg['Ly']}", flush=True)
    st = random_state(M, N, L, seed=M * 3 + N, amp=1e-2, decay=0.05)
    keep = kept_mask(M, N)
    o = _orc(cfg)
    ref_b, ref_i = o.inverse_B(*st), o.implicit(*st)
    o.set_state(*st)
    o.step(5)
    ref = o.get_state()
    o.close()
    with _gpu(cfg, "f64") as G:
        eb = _cmp(G.calculate_inverse(*st), ref_b, keep)
        ei = _cmp(G.calculate_implicit(*st), ref_i, keep)
        G.set_state(*st)
        G.step_time(5)
        es = _cmp(G.get_state(), ref, keep)
    print("inverse B", eb, "implicit", ei, "5 steps", es)
    assert max(eb) < 1e-14 and max(ei) < 1e-13 and max(es) < 1e-10, (eb, ei, es)
    with _gpu(cfg, "f32") as G:
        G.set_state(*st)
        G.step_time(5)
        e32 = _cmp(G.get_state(), ref, keep)
    print("fp32 5 steps", e32)
    assert max(e32[:3]) < 1e-4, e32
