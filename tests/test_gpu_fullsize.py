"""Parity at the full 512 x 512 x 32 workload through size-independent properties: states
band-limited to a small grid's wavenumbers give the same operators, nonlinear term, step and
point velocity on any grid that keeps those wavenumbers, so the oracle at 32^2 x 32 is the
reference for the 512^2 x 32 GPU path.

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


# The nonlinear term of a state made of two horizontal modes lives on the wavenumbers k1 +- k2,
# 2 k1, 2 k2 and 0, all of which both grids keep, and the collocation form is exact for such
# band-limited input: the 512^2 x 32 term must equal the oracle's at 32^2 x 32 on every
# wavenumber of the small grid and vanish elsewhere.  Tolerances relative to the largest
# coefficient of each field: FP64 1e-12, FP32 2e-5, BF16X3 2e-4 (as the operator tests); the
# "elsewhere" bound is 10 x that in FP32 and BF16X3: the rounding of the 512-point transforms
# spreads over every wavenumber, and the f and g equations weight it by wavenumber factors up
# to |k| ~ 270 against |k| ~ 5 of the signal (measured: f 1.9e-5, g 8.1e-5 in FP32).
TWO = [(1, 1), (-2, 3)]  # (ky, kx), both kx > 0


def _two_mode_state(M, N, seed=5):
    mox = M // 2 + 1
    fields = []
    for c in range(3):
        a = np.zeros((L, N, mox), np.complex128)
        z = gaussian_complex(seed + c, L * len(TWO)).reshape(len(TWO), L) * 1e-2
        for i, (ky, kx) in enumerate(TWO):
            a[:, ky % N, kx] = z[i]
        fields.append(a)
    mean = gaussian_complex(seed + 3, 2 * L).reshape(2, L).real * 1e-2
    return fields + [mean[0].astype(np.complex128), mean[1].astype(np.complex128)]


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 2e-5), ("bf16x3", 2e-4)])
def test_two_mode_nonlinear_term_matches_the_oracle_on_a_small_grid(precision, tol):
    from oracle import oracle as O
    from sdc_amd import Stepper
    print(f"case: two-mode nonlinear term 512x512x32 {precision} vs oracle 32x32x32", flush=True)
    O.build()
    o = O.Stepper(32, 32, L, *PHYS, nthreads=16)
    ref = o.nonlinear(*_two_mode_state(32, 32))
    o.close()
    with Stepper(512, 512, L, *PHYS, precision=precision) as S:
        out = S.calculate_nonlinear(*_two_mode_state(512, 512))
    keep32 = kept_mask(32, 32)
    rows = np.arange(32)
    ky = np.where(rows < 16, rows, rows - 32)
    errs, leaks = [], []
    for c in range(3):
        big = out[c][:, ky % 512, :17]  # the small grid's wavenumbers inside the large grid
        scale = np.abs(ref[c]).max()
        assert scale > 0
        errs.append(np.abs(big[:, keep32] - ref[c][:, keep32]).max() / scale)
        rest = out[c].copy()
        rest[:, (ky % 512)[:, None], np.arange(17)[None, :]] = 0
        leaks.append(np.abs(rest).max() / scale)
    for c in (3, 4):
        errs.append(np.abs(out[c] - ref[c]).max() / max(np.abs(ref[c]).max(), 1e-300))
    print("errors (theta, f, g, F, G)", errs, "elsewhere", leaks)
    assert max(errs) < tol, errs
    assert max(leaks) < (tol if precision == "f64" else 10 * tol), leaks


# The spectral velocity at arbitrary points (the tracer evaluation, matrix-core kernels with
# L = 32) of a state band-limited to the small grid's kept modes is the same Fourier sum on
# either grid: rbc_velocity_at at 512^2 x 32 against oracle/velocity.py on the 32^2 x 32 state,
# 5,000 random points.  Tolerances relative to the largest velocity component: 2e-5 for FP64
# and FP32 states (the point evaluation runs in FP32 for both: float positions and velocities,
# as the tracers keep them; measured 2e-6), 2e-4 for BF16X3.
@pytest.mark.parametrize("precision,tol", [("f64", 2e-5), ("f32", 2e-5), ("bf16x3", 2e-4)])
def test_point_velocity_matches_the_oracle_on_a_small_grid(precision, tol):
    from oracle import velocity as V
    from sdc_amd import Stepper
    print(f"case: point velocity 512x512x32 {precision} vs oracle 32x32x32", flush=True)
    keep = kept_mask(32, 32)
    rows, cols = np.nonzero(keep)
    pick = np.random.default_rng(8).choice(rows.size, 24, replace=False)
    sel = [(int(rows[i]), int(cols[i])) for i in pick if cols[i] > 0]  # kx > 0: no mirrored partner
    ky = [r if r < 16 else r - 32 for r, _ in sel]

    def state(M, N):
        mox = M // 2 + 1
        out = []
        for c in range(3):
            a = np.zeros((L, N, mox), np.complex128)
            z = gaussian_complex(40 + c, L * len(sel)).reshape(len(sel), L) * 1e-2
            for i, (k, (_, kx)) in enumerate(zip(ky, sel)):
                a[:, k % N, kx] = z[i]
            out.append(a)
        mean = gaussian_complex(43, 2 * L).reshape(2, L).real * 1e-2
        return out + [mean[0].astype(np.complex128), mean[1].astype(np.complex128)]

    rng = np.random.default_rng(9)
    n = 5000
    x = rng.uniform(0, PHYS[0], n).astype(np.float32)
    y = rng.uniform(0, PHYS[1], n).astype(np.float32)
    z = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    ref = V.velocity_points(state(32, 32), 32, 32, PHYS[0], PHYS[1], x.astype(np.float64), y.astype(np.float64),
                            z.astype(np.float64))
    with Stepper(512, 512, L, *PHYS, precision=precision) as S:
        S.set_state(*state(512, 512))
        got = S.velocity_at(x, y, z)
    errs = [np.abs(a - b).max() / np.abs(b).max() for a, b in zip(got, ref)]
    print("u, v, w errors", errs)
    assert max(errs) < tol, errs


# The linear operators act mode by mode, and one step of a state band-limited to |kx|, |ky| <= 5
# forms products up to |k| <= 10, which the 32^2 grid keeps: inverse B, the implicit operator
# and one (Euler) step at 512^2 x 32 must equal the oracle's at 32^2 x 32 on every wavenumber of
# the small grid (the 512^2 factor records group many more wavenumbers per |k|^2: this pins
# their grouping at the full size) and vanish elsewhere.  Tolerances relative to each field's
# largest coefficient: operators FP64 1e-13, FP32 and BF16X3 2e-6 (factors rounded to FP32);
# one step FP64 1e-10, FP32 2e-5, BF16X3 1e-4.
def _band_limited_32(seed=12):
    from tests.helpers import random_state
    st = random_state(32, 32, L, seed=seed, amp=1e-2, decay=0.05)
    rows = np.arange(32)
    ky = np.where(rows < 16, rows, rows - 32)
    cut = (np.abs(ky)[:, None] > 5) | (np.arange(17)[None, :] > 5)
    for a in st[:3]:
        a[:, cut] = 0
    return st


def _to_512(st32):
    rows = np.arange(32)
    ky = np.where(rows < 16, rows, rows - 32)
    out = []
    for a in st32[:3]:
        b = np.zeros((L, 512, 257), np.complex128)
        b[:, ky % 512, :17] = a
        out.append(b)
    return out + [st32[3].copy(), st32[4].copy()]


def _cmp_512_32(out, ref):
    rows = np.arange(32)
    ky = np.where(rows < 16, rows, rows - 32)
    keep = kept_mask(32, 32)
    errs, leaks = [], []
    for c in range(3):
        scale = np.abs(ref[c]).max()
        big = out[c][:, ky % 512, :17]
        errs.append(np.abs(big[:, keep] - ref[c][:, keep]).max() / scale)
        rest = out[c].copy()
        rest[:, (ky % 512)[:, None], np.arange(17)[None, :]] = 0
        leaks.append(np.abs(rest).max() / scale)
    for c in (3, 4):
        errs.append(np.abs(out[c] - ref[c]).max() / max(np.abs(ref[c]).max(), 1e-300))
    return errs, leaks


@pytest.mark.parametrize("precision,tol_op,tol_step", [("f64", 1e-13, 1e-10), ("f32", 2e-6, 2e-5),
                                                       ("bf16x3", 2e-6, 1e-4)])
def test_operators_and_one_step_match_the_oracle_on_a_small_grid(precision, tol_op, tol_step):
    from oracle import oracle as O
    from sdc_amd import Stepper
    print(f"case: operators and one step 512x512x32 {precision} vs oracle 32x32x32", flush=True)
    O.build()
    st32 = _band_limited_32()
    o = O.Stepper(32, 32, L, *PHYS, nthreads=16)
    ref_b, ref_i = o.inverse_B(*st32), o.implicit(*st32)
    o.set_state(*st32)
    o.step(1)
    ref_s = o.get_state()
    o.close()
    st = _to_512(st32)
    with Stepper(512, 512, L, *PHYS, precision=precision) as S:
        eb, lb = _cmp_512_32(S.calculate_inverse(*st), ref_b)
        ei, li = _cmp_512_32(S.calculate_implicit(*st), ref_i)
        S.set_state(*st)
        S.step_time(1)
        es, ls = _cmp_512_32(S.get_state(), ref_s)
    print("inverse B", eb, lb, "implicit", ei, li, "one step", es, ls)
    assert max(eb + ei) < tol_op and max(lb + li) == 0, (eb, ei, lb, li)
    assert max(es) < tol_step, es
    assert max(ls) < (tol_step if precision == "f64" else 10 * tol_step), ls
