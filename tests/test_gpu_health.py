"""Non-finite state detection (≙ find_NaN, util/util.cu:83-105): a step that ends with a NaN
in the state is reported as RBC_ERR_NUMERIC with its step index by rbc_synchronize,
rbc_get_state and rbc_nusselt; the next rbc_step refuses until the state is replaced.  The
NaN enters through rbc_set_state (host data), so every kernel runs on ordinary memory."""
import numpy as np
import pytest

from tests.helpers import random_state

pytestmark = pytest.mark.gpu

CFG = (32, 32, 8, 4.032, 4.032, 5000.0, 1.0, 1e-3)


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16x3"])
def test_nan_in_one_theta_coefficient_is_reported(precision):
    print("case: NaN injection", precision, flush=True)
    from sdc_amd import RbcNumericError, Stepper
    st = random_state(32, 32, 8, seed=3, amp=1e-2)
    with Stepper(*CFG, precision=precision) as S:
        S.set_state(*st)
        S.step_time(3)
        S.synchronize()  # finite so far
        bad = [a.copy() for a in S.get_state()]
        bad[0][2, 3, 5] = np.nan  # one theta coefficient (m = 2, ky row 3, kx 5)
        S.set_state(*bad)
        S.step_time(2)
        with pytest.raises(RbcNumericError, match="step 4 "):
            S.synchronize()
        with pytest.raises(RbcNumericError):
            S.get_state()
        with pytest.raises(RbcNumericError):
            S.nusselt()
        with pytest.raises(RbcNumericError, match="refusing"):
            S.step_time(1)
        # a finite state clears the flag; the AB2 history still holds the NaN steps' terms, so
        # the restart is an Euler start, and stepping resumes
        S.set_state(*st)
        S.reset_history()
        S.step_time(2)
        S.synchronize()
        out = S.get_state()
        assert all(np.isfinite(a).all() for a in out)


def test_finite_runs_stay_clean_in_ensembles():
    """The health words of ensemble members advance inside the graph; a healthy ensemble
    reports OK from every member."""
    print("case: ensemble health", flush=True)
    from sdc_amd import Ensemble
    cfgs = [dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=2e4, Pr=0.7, dt=5e-4, precision="f32")] * 2
    with Ensemble(cfgs) as E:
        for i, m in enumerate(E.members):
            m.set_state(*random_state(64, 64, 8, seed=40 + i, amp=1e-3))
        E.step(12, steps_per_launch=5)
        for m in E.members:
            m.synchronize()
            assert all(np.isfinite(a).all() for a in m.get_state())
