"""Ensembles (rbc_ensemble_*): graph-replayed batches of uneven runs equal the same runs
stepped alone, bit for bit, in both graph and eager modes; the FP64 member matches the oracle
and an FP32 member matches its FP64 twin."""
import numpy as np
import pytest

from oracle import oracle as O
from sdc_amd.initial import kept_mask, random_theta

pytestmark = pytest.mark.gpu

CONFIGS = [
    dict(M=32, N=32, L=8, Lx=4.032, Ly=4.032, Ra=5000.0, Pr=1.0, dt=1e-3, precision="f64"),
    dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=2e4, Pr=0.7, dt=5e-4, precision="f32"),
    dict(M=64, N=32, L=16, Lx=4.0, Ly=2.0, Ra=1e5, Pr=0.7, dt=2e-4, precision="f32"),
    dict(M=128, N=128, L=24, Lx=4.0, Ly=4.0, Ra=1e6, Pr=0.7, dt=5e-5, precision="f32"),
]
STEPS = 12


def _states():
    return [random_theta(c["M"], c["N"], c["L"], seed=1000 + i, amp=1e-3) for i, c in enumerate(CONFIGS)]


def _alone(cfg, st, precision=None):
    from sdc_amd import Stepper
    with Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                 precision=precision or cfg["precision"]) as S:
        S.set_state(*st)
        S.step_time(STEPS)
        return S.get_state()


@pytest.fixture(scope="module")
def singles():
    O.build()
    return [_alone(c, s) for c, s in zip(CONFIGS, _states())]


@pytest.mark.parametrize("spl", [5, 0])
def test_ensemble_equals_single_runs(singles, spl):
    print("case: ensemble steps_per_launch", spl, flush=True)
    from sdc_amd import Ensemble
    with Ensemble(CONFIGS) as E:
        rep = E.tune()
        print("tuned (y-FFT columns, row threads, ms/step):", rep.tolist())
        assert rep.shape == (len(CONFIGS), 3) and np.all(rep[:, 2] > 0)
        for m, st in zip(E.members, _states()):
            m.set_state(*st)
        E.step(STEPS, steps_per_launch=spl)  # Euler start, 2 replays of 5 steps, 1 eager step
        got = [m.get_state() for m in E.members]
    for i, (g, r) in enumerate(zip(got, singles)):
        for a, b in zip(g, r):
            assert np.array_equal(a, b), f"member {i}"


def test_ensemble_fp64_member_vs_oracle(singles):
    c = CONFIGS[0]
    o = O.Stepper(c["M"], c["N"], c["L"], c["Lx"], c["Ly"], c["Ra"], c["Pr"], c["dt"], nthreads=8)
    o.set_state(*_states()[0])
    o.step(STEPS)
    ref = o.get_state()
    o.close()
    keep = kept_mask(c["M"], c["N"])
    th = singles[0][0]
    err = np.abs(th[:, keep] - ref[0][:, keep]).max() / np.abs(ref[0][:, keep]).max()
    assert err < 1e-10, err


def test_ensemble_fp32_member_vs_fp64_twin(singles):
    c = CONFIGS[3]
    ref = _alone(c, _states()[3], precision="f64")
    keep = kept_mask(c["M"], c["N"])
    err = np.abs(singles[3][0][:, keep] - ref[0][:, keep]).max() / np.abs(ref[0][:, keep]).max()
    print("128^2 x 24 FP32 vs FP64 after", STEPS, "steps:", err)
    assert err < 1e-4, err


@pytest.mark.parametrize("spl", [5, 0])
def test_ensemble_mixed_history_steps_every_member_n_times(spl):
    """A member reset to an Euler start beside members that keep their AB2 history: one
    rbc_ensemble_step(n) advances every member by exactly n steps (the reset member by an
    Euler step then AB2, the others by AB2 throughout), equal to solo contexts stepped the
    same way."""
    print("case: ensemble mixed history, steps_per_launch", spl, flush=True)
    from sdc_amd import Ensemble, Stepper
    cfgs = CONFIGS[1:3]
    states = [random_theta(c["M"], c["N"], c["L"], seed=2000 + i, amp=1e-3) for i, c in enumerate(cfgs)]
    first, second = 3, 11
    solo = []
    for i, (c, st) in enumerate(zip(cfgs, states)):
        with Stepper(c["M"], c["N"], c["L"], c["Lx"], c["Ly"], c["Ra"], c["Pr"], c["dt"],
                     precision=c["precision"]) as S:
            S.set_state(*st)
            S.step_time(first)
            if i == 0:
                S.reset_history()
            S.step_time(second)
            solo.append(S.get_state())
    with Ensemble(cfgs) as E:
        for m, st in zip(E.members, states):
            m.set_state(*st)
        E.step(first, steps_per_launch=spl)
        E.members[0].reset_history()
        E.step(second, steps_per_launch=spl)
        got = [m.get_state() for m in E.members]
    for i, (g, r) in enumerate(zip(got, solo)):
        for a, b in zip(g, r):
            assert np.array_equal(a, b), f"member {i}"
