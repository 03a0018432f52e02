"""Learned closure on the GPU (rbc_closure_*) against the FP64 restatement in
oracle/closure.py: the network's gradient, prediction and Adam step, the training data
pipeline (inputs and residual targets at the coarse points), the correction inside a coarse
step, the fallback to the plain method, and the gain of a trained network over no closure on
held-out data.

Tolerances: gradient and prediction 1e-4 relative (BF16 weights are shared, activations
are FP32 on the GPU); point data 2e-5 of the field maximum (FP32 synthesis)."""
import numpy as np
import pytest

from oracle import closure as C
from oracle import oracle as O
from sdc_amd.initial import random_theta
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

FINE = dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=3e4, Pr=0.7, dt=2e-4)
COARSE = dict(M=32, N=32, L=4, Lx=4.0, Ly=4.0, Ra=3e4, Pr=0.7, dt=2e-4)


def _stepper(cfg, precision="f64"):
    from sdc_amd import Stepper
    return Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                   precision=precision)


def _net(seed=0, n=3000):
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal(C.NW) * 0.2).astype(np.float32)
    norm = (rng.standard_normal(C.NIN).astype(np.float32) * 0.1, (1.0 + rng.random(C.NIN)).astype(np.float32),
            (0.5 + rng.random(C.NOUT)).astype(np.float32))
    X = rng.standard_normal((n, C.NIN)).astype(np.float32)
    Y = rng.standard_normal((n, C.NOUT)).astype(np.float32)
    return w, norm, X, Y


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def test_gradient_and_prediction_vs_oracle():
    print("case: closure gradient", flush=True)
    from sdc_amd import Closure
    w, norm, X, Y = _net()
    with _stepper(COARSE, "f32") as S:
        cl = Closure(S, capacity_points=0)
        cl.set_weights(w)
        cl.set_normalization(*norm)
        g, loss = cl.gradient(X, Y)
        P = cl.predict(X[:500])
        cl.close()
    wb = C.bf16_round(w).astype(np.float64)  # the kernels use the BF16 copies
    lr, gr = C.loss_and_grad(wb, norm, X, Y)
    assert abs(loss - lr) < 1e-5 * lr
    assert np.abs(g - gr).max() < 1e-4 * np.abs(gr).max(), np.abs(g - gr).max() / np.abs(gr).max()
    pr = C.forward(wb, norm, X[:500])
    assert np.abs(P - pr).max() < 1e-4 * np.abs(pr).max()


def test_adam_step_vs_oracle():
    """All dataset points identical: one minibatch step is one Adam step on that point's
    gradient, whatever points are drawn."""
    print("case: closure adam", flush=True)
    from sdc_amd import Closure
    w, norm, X, Y = _net(1)
    Xs = np.repeat(X[:1], 1000, axis=0)
    Ys = np.repeat(Y[:1], 1000, axis=0)
    with _stepper(COARSE, "f32") as S:
        cl = Closure(S, capacity_points=1000)
        cl.set_weights(w)
        cl.set_normalization(*norm)
        cl.load_dataset(Xs, Ys)
        loss = cl.train(1, batch=256, lr=1e-3, seed=7)
        w1 = cl.weights()
        cl.close()
    wb = C.bf16_round(w).astype(np.float64)
    l0, g = C.loss_and_grad(wb, norm, X[:1], Y[:1])
    assert abs(loss[0] - l0) < 1e-5 * l0
    wr, _, _ = C.adam(w.astype(np.float64), g, 0 * g, 0 * g, 1, 1e-3)
    sel = np.abs(g) > 1e-5 * np.abs(g).max()
    assert np.abs((w1 - w)[sel] - (wr - w)[sel]).max() < 2e-6, np.abs((w1 - w)[sel] - (wr - w)[sel]).max()


def test_training_reduces_loss_on_a_learnable_map():
    print("case: closure training", flush=True)
    from sdc_amd import Closure
    rng = np.random.default_rng(2)
    n = 40000
    X = rng.standard_normal((n, C.NIN)).astype(np.float32)
    Y = np.stack([np.sin(X[:, 0]) * X[:, 1], np.tanh(X[:, 2] + X[:, 3]), 0.5 * X[:, 4] * X[:, 5]],
                 axis=1).astype(np.float32)
    with _stepper(COARSE, "f32") as S:
        cl = Closure(S, capacity_points=n)
        cl.load_dataset(X, Y)
        cl.fit_normalization(30000)
        loss = cl.train(400, batch=4096, lr=3e-3, seed=1, first=0, n=30000)
        mse, zero = cl.evaluate(30000, 10000)
        cl.close()
    print("loss first/last", loss[0], loss[-1], "held-out mse", mse, "zero", zero)
    assert loss[-1] < 0.2 * loss[0]
    assert mse < 0.2 * zero


@pytest.mark.parametrize("fm,cm", [(64, 32), (128, 64)])
def test_point_data_vs_oracle(fm, cm):
    """Inputs and residual targets of a sample (fine fm^2 x 8 -> coarse cm^2 x 4, FP64
    contexts; cm = 64 runs the group-transform synthesis) against the oracle's nonlinear terms,
    truncation and synthesis."""
    print(f"case: closure sample {fm} -> {cm}", flush=True)
    from sdc_amd import Closure
    fine, coarse = dict(FINE, M=fm, N=fm), dict(COARSE, M=cm, N=cm)
    st = random_state(fm, fm, fine["L"], seed=9, amp=1e-1)
    with _stepper(fine) as F, _stepper(coarse) as Cc:
        F.set_state(*st)
        cl = Closure(Cc, nz=8, capacity_points=cm * cm * 8)
        X, Y = cl.points(F)
        nz = cl.nz
        cl.close()
    of = O.Stepper(fm, fm, fine["L"], fine["Lx"], fine["Ly"], fine["Ra"], fine["Pr"], fine["dt"], nthreads=8)
    oc = O.Stepper(cm, cm, coarse["L"], coarse["Lx"], coarse["Ly"], coarse["Ra"], coarse["Pr"], coarse["dt"],
                   nthreads=8)
    Nf = of.nonlinear(*st)
    stc = [C.truncate(a, cm, cm, 4) for a in st]
    Nc = oc.nonlinear(*stc)
    keep = C.kept_mask(cm, cm)  # the oracle also returns the removed band, which no step sees
    R = [C.truncate(Nf[i], cm, cm, 4) - Nc[i] * keep for i in range(3)]
    z, _ = O.gauss_legendre(nz)
    Xr = C.features(stc, cm, cm, coarse["Lx"], coarse["Ly"], z)
    Yr = C.residual_points(R, cm, cm, z)
    of.close()
    oc.close()
    ex = [float(np.abs(X[:, j] - Xr[:, j]).max() / max(np.abs(Xr[:, j]).max(), 1e-30)) for j in range(C.NIN)]
    ey = [float(np.abs(Y[:, j] - Yr[:, j]).max() / np.abs(Yr[:, j]).max()) for j in range(C.NOUT)]
    print("input errors", ex, "target errors", ey)
    assert max(ex) <= 2e-5 and max(ey) <= 2e-5


@pytest.mark.parametrize("cm", [32, 64])
def test_correction_inside_the_step(cm):
    """One corrected Euler step differs from the plain one by Implicit(dt B^-1 dN), dN the
    projection of the network's prediction at the points (cm = 64: group-transform analysis)."""
    print(f"case: closure correction {cm}", flush=True)
    from sdc_amd import Closure
    coarse = dict(COARSE, M=cm, N=cm)
    st = random_state(cm, cm, coarse["L"], seed=12, amp=1e-1)
    w, norm, _, _ = _net(3)
    with _stepper(coarse) as A, _stepper(coarse) as B:
        A.set_state(*st)
        B.set_state(*st)
        cl = Closure(A, nz=8, capacity_points=cm * cm * 8)
        cl.set_weights(w)
        cl.set_normalization(*norm)
        X, _ = cl.points(A)  # inputs of A's own state
        P = cl.predict(X)
        cl.enable(True, max_ratio=1e30)
        A.step_time(1)
        cl.enable(False)
        B.step_time(1)
        xa, xb = A.get_state(), B.get_state()
        z, wz = O.gauss_legendre(cl.nz)
        dN = C.project_points(P.astype(np.float64), cm, cm, 4, z, wz)
        zeros = np.zeros(4, complex)
        r = B.calculate_inverse(dN[0], dN[1], dN[2], zeros, zeros)
        e = B.calculate_implicit(*[coarse["dt"] * a for a in r[:3]], zeros, zeros)
        cl.close()
    for i in range(3):
        d = xa[i] - xb[i]
        assert np.abs(d - e[i]).max() < 1e-5 * np.abs(e[i]).max(), i


def test_correction_interval_reuses_the_last_correction():
    """rbc_closure_interval(k): the network runs on every k-th corrected step and the steps
    between add the last correction again.  With W1 = 0 the network's output does not depend on
    the state, so re-evaluating and reusing give the same correction and k = 1, 2, 3 must agree
    bit for bit; with a state-dependent network k = 2 must differ from k = 1 after the second
    step (the reused correction is older) and agree after the first."""
    print("case: closure interval", flush=True)
    from sdc_amd import Closure
    st = random_state(COARSE["M"], COARSE["N"], COARSE["L"], seed=14, amp=1e-1)
    w, norm, _, _ = _net(5)
    w_const = w.copy()
    w_const[C.OFF_W1:C.OFF_B1] = 0.0

    def run(weights, every, nsteps):
        with _stepper(COARSE, "f32") as A:
            A.set_state(*st)
            cl = Closure(A, nz=8, capacity_points=0)
            cl.set_weights(weights)
            cl.set_normalization(*norm)
            cl.set_interval(every)
            cl.enable(True, max_ratio=1e30)
            cl.step(nsteps, check_every=nsteps)
            out = A.get_state()
            cl.close()
        return out

    ref = run(w_const, 1, 4)
    for k in (2, 3):
        got = run(w_const, k, 4)
        assert all(np.array_equal(a, b) for a, b in zip(got, ref)), k
    one1, one2 = run(w, 1, 1), run(w, 2, 1)
    assert all(np.array_equal(a, b) for a, b in zip(one1, one2))
    two1, two2 = run(w, 1, 2), run(w, 2, 2)
    assert max(float(np.abs(a - b).max()) for a, b in zip(two1, two2)) > 0.0
    assert all(np.isfinite(a).all() for a in two2)


def test_damping_scales_and_filters_the_correction():
    """rbc_closure_damping(gain, kfrac): gain 0 gives the plain step bit for bit; one corrected
    Euler step is linear in the added correction, so the change a gain of 0.5 makes is half the
    change of gain 1 (to FP32 rounding of the update); with kfrac = 0.5 the change vanishes on
    every mode with |k| > 0.5 k_max and equals the undamped change on the others."""
    print("case: closure damping", flush=True)
    from sdc_amd import Closure
    from sdc_amd.initial import kept_mask
    st = random_state(COARSE["M"], COARSE["N"], COARSE["L"], seed=15, amp=1e-1)
    w, norm, _, _ = _net(6)

    def run(gain, kfrac, closure=True):
        with _stepper(COARSE, "f64") as A:
            A.set_state(*st)
            cl = Closure(A, nz=8, capacity_points=0)
            cl.set_weights(w)
            cl.set_normalization(*norm)
            cl.set_damping(gain, kfrac)
            cl.enable(closure, max_ratio=1e30)
            if closure:
                cl.step(1, check_every=1)
            else:
                A.step_time(1)
            out = A.get_state()
            cl.close()
        return out

    plain = run(1.0, 1.0, closure=False)
    zero = run(0.0, 1.0)
    assert all(np.array_equal(a, b) for a, b in zip(zero, plain))
    full, halfg, low = run(1.0, 1.0), run(0.5, 1.0), run(1.0, 0.5)
    d_full = [a - b for a, b in zip(full[:3], plain[:3])]
    d_half = [a - b for a, b in zip(halfg[:3], plain[:3])]
    d_low = [a - b for a, b in zip(low[:3], plain[:3])]
    scale = max(float(np.abs(d).max()) for d in d_full)
    assert scale > 0
    assert max(float(np.abs(h - 0.5 * f).max()) for h, f in zip(d_half, d_full)) < 1e-9 * max(scale, 1.0)
    # |k| per (ky row, kx column) of the reference layout
    M, N = COARSE["M"], COARSE["N"]
    kx = 2 * np.pi / COARSE["Lx"] * np.arange(M // 2 + 1)
    ky = 2 * np.pi / COARSE["Ly"] * np.fft.fftfreq(N, 1.0 / N)
    k = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    keep = kept_mask(M, N)
    high = keep & (k > 0.5 * k[keep].max() * (1 + 1e-12))
    lowm = keep & (k < 0.5 * k[keep].max() * (1 - 1e-12))
    assert high.any() and lowm.any()
    for dl, df in zip(d_low, d_full):
        assert np.array_equal(dl[:, high], np.zeros_like(dl[:, high]))
        assert np.abs(dl[:, lowm] - df[:, lowm]).max() < 1e-9 * max(scale, 1.0)


def test_fallback_to_the_plain_step():
    print("case: closure fallback", flush=True)
    from sdc_amd import Closure
    st = random_state(COARSE["M"], COARSE["N"], COARSE["L"], seed=13, amp=1e-1)
    w, norm, _, _ = _net(4)
    with _stepper(COARSE) as A, _stepper(COARSE) as B:
        B.set_state(*st)
        B.step_time(6)
        ref = B.get_state()
        for bad in ("nan", "ratio"):
            A.set_state(*st)
            A.reset_history()
            cl = Closure(A, nz=8, capacity_points=0)
            cl.set_normalization(*norm)
            if bad == "nan":
                cl.set_weights(np.full(C.NW, np.nan, np.float32))
                cl.enable(True, max_ratio=10.0)
            else:
                cl.set_weights(w)
                cl.enable(True, max_ratio=1e-12)
            cl.step(6, check_every=3)
            active, fell_back, ratio = cl.status()
            got = A.get_state()
            diff = [float(np.abs(a - b).max()) for a, b in zip(got, ref)]
            print(bad, "active", active, "fallback", fell_back, "ratio", ratio, "max diff vs plain", diff)
            assert fell_back and not active, (bad, ratio)
            assert max(diff) == 0.0, bad
            cl.close()


def test_trained_closure_beats_no_closure_on_held_out_data():
    """A-priori gain: train on two snapshots of a developing 128^2 x 16 run (coarse grid
    64^2 x 8) and compare the held-out error of the network with predicting no correction."""
    print("case: closure a-priori gain", flush=True)
    from sdc_amd import Closure
    fine = dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=4e-5)
    coarse = dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=4e-5)
    with _stepper(fine, "f32") as F, _stepper(coarse, "f32") as Cc:
        F.set_state(*random_theta(128, 128, 16, seed=21, amp=1e-3))
        for _ in range(10):
            F.step_time(500)
            print("spin-up Nu", F.nusselt(), flush=True)
        cl = Closure(Cc, nz=16, capacity_points=3 * 64 * 64 * 16)
        for _ in range(3):
            cl.sample(F)
            F.step_time(100)
        npts = cl.npts
        cl.fit_normalization(2 * npts)
        loss = cl.train(600, batch=8192, lr=2e-3, seed=3, first=0, n=2 * npts)
        mse, zero = cl.evaluate(2 * npts, npts)
        nu = F.nusselt()
        cl.close()
    print("Nu", nu, "loss first/last", loss[0], loss[-1], "held-out mse", mse, "no-closure mse", zero)
    assert mse < 0.9 * zero


def test_workflow_tolerance_fallback():
    """The experiment driver keeps the plain coarse result when the closure's Nu misses the
    fine run's by more than the tolerance (tolerance 0 here), and reports both."""
    print("case: closure workflow fallback", flush=True)
    from sdc_amd import closure
    fine = dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=3e4, Pr=0.7, dt=2e-4)
    coarse = dict(M=32, N=32, L=4, Lx=4.0, Ly=4.0, Ra=3e4, Pr=0.7, dt=4e-4)
    r = closure.run(fine, coarse, T=0.04, spinup_steps=200, samples=3, train_steps=20, batch=1024, tol=0.0,
                    every=50)
    assert r["fallback_to_plain"] is True
    assert r["result_nu"] == r["plain_coarse"]["nu"]
    assert np.isfinite(r["closure_coarse"]["nu"]) and np.isfinite(r["nu_fine"])
    assert r["training"]["points"] == 3 * 32 * 32 * 8
