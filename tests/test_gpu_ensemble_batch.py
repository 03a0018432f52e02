"""Ensembles with several members per size class: the members of a class advance as one batch
(one launch per stage for the class, blockIdx.z = member) inside the captured graph.  Every
member must equal the same run stepped alone, bit for bit: members of one class differ in Ra,
dt and box size (so in their factor tables and |k|^2 chunking), and the classes cover the FP32
and BF16X3 matrix-core paths, a non-square grid, and members that are not batched (a class of
one, a class of large members, an FP64 member)."""
import numpy as np
import pytest

from sdc_amd.initial import random_theta

pytestmark = pytest.mark.gpu

CONFIGS = [
    dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=2e4, Pr=0.7, dt=5e-4, precision="f32"),
    dict(M=64, N=64, L=8, Lx=3.5, Ly=4.0, Ra=5e4, Pr=0.7, dt=4e-4, precision="f32"),
    dict(M=64, N=64, L=8, Lx=4.0, Ly=4.0, Ra=1e5, Pr=1.0, dt=3e-4, precision="f32"),
    dict(M=128, N=64, L=16, Lx=4.0, Ly=2.0, Ra=1e5, Pr=0.7, dt=2e-4, precision="f32"),
    dict(M=128, N=64, L=16, Lx=4.0, Ly=2.0, Ra=3e5, Pr=0.7, dt=1e-4, precision="f32"),
    dict(M=64, N=64, L=16, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=2e-4, precision="bf16x3"),
    dict(M=64, N=64, L=16, Lx=4.0, Ly=4.0, Ra=4e5, Pr=0.7, dt=1e-4, precision="bf16x3"),
    dict(M=64, N=64, L=16, Lx=4.0, Ly=4.0, Ra=5e4, Pr=0.7, dt=3e-4, precision="f32"),
    dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=1e5, Pr=0.7, dt=2e-4, precision="f32"),
    dict(M=128, N=128, L=16, Lx=4.0, Ly=4.0, Ra=3e5, Pr=0.7, dt=1e-4, precision="f32"),
    dict(M=32, N=32, L=8, Lx=4.032, Ly=4.032, Ra=5000.0, Pr=1.0, dt=1e-3, precision="f64"),
]
# three classes of 3, 2 and 2 members; the 64^2 x 16 FP32 member is a class of one, the 128^2
# x 16 members have more stage data per member than the batches take (they fill the device
# alone), the FP64 member is off the batched path
BATCHED = 7
STEPS = 12


def _states():
    return [random_theta(c["M"], c["N"], c["L"], seed=2000 + i, amp=1e-3) for i, c in enumerate(CONFIGS)]


def _alone(cfg, st):
    from sdc_amd import Stepper
    with Stepper(cfg["M"], cfg["N"], cfg["L"], cfg["Lx"], cfg["Ly"], cfg["Ra"], cfg["Pr"], cfg["dt"],
                 precision=cfg["precision"]) as S:
        S.set_state(*st)
        S.step_time(STEPS)
        return S.get_state()


@pytest.mark.parametrize("tune", [False, True])
def test_batched_ensemble_equals_single_runs(tune):
    print("case: batched ensemble, tuned" if tune else "case: batched ensemble, default shapes", flush=True)
    from sdc_amd import Ensemble
    singles = [_alone(c, s) for c, s in zip(CONFIGS, _states())]
    with Ensemble(CONFIGS) as E:
        if tune:
            print("tuned:", E.tune().tolist())
        for m, st in zip(E.members, _states()):
            m.set_state(*st)
        E.step(STEPS, steps_per_launch=5)  # Euler start, 2 replays of 5 batched steps, 1 eager step
        assert E.batched_members == BATCHED, E.batched_members
        got = [m.get_state() for m in E.members]
    for i, (g, r) in enumerate(zip(got, singles)):
        assert all(np.all(np.isfinite(a)) for a in g), f"member {i}"
        for a, b in zip(g, r):
            assert np.array_equal(a, b), f"member {i} ({CONFIGS[i]['M']}x{CONFIGS[i]['N']}x{CONFIGS[i]['L']})"


def test_batching_off_steps_every_member_alone():
    print("case: ensemble with batching off", flush=True)
    from sdc_amd import Ensemble
    sts = _states()[:3]
    with Ensemble(CONFIGS[:3]) as E:
        E.set_batching(False)
        for m, st in zip(E.members, sts):
            m.set_state(*st)
        E.step(6, steps_per_launch=5)
        assert E.batched_members == 0
        off = [m.get_state() for m in E.members]
    with Ensemble(CONFIGS[:3]) as E:
        for m, st in zip(E.members, sts):
            m.set_state(*st)
        E.step(6, steps_per_launch=5)
        assert E.batched_members == 3
        on = [m.get_state() for m in E.members]
    for a, b in zip(off, on):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
