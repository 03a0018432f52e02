"""The no-allocation entry point rbc_velocity_at_device_noalloc (partial sums in the library's
static per-GPU scratch) against rbc_velocity_at_device (partial sums in the context's scratch):
the same plan and kernels, so the results are equal bit for bit, on two calls in a row with the
second smaller than the first; and a point count beyond the scratch is rejected before any
GPU work."""
import numpy as np
import pytest

from tests.helpers import random_state
from tests.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,L", [("f32", 16), ("f64", 8)])
def test_noalloc_velocity_equals_allocating_entry_point(precision, L):
    print("case: no-alloc velocity", precision, L, flush=True)
    from sdc_amd import RbcError, Stepper
    cfg = (64, 64, L, 4.0, 3.0, 2e4, 0.7, 1e-3)
    st = random_state(64, 64, L, seed=91, amp=5e-2, decay=0.1)
    rng = np.random.default_rng(12)
    with Stepper(*cfg, precision=precision) as S:
        S.set_state(*st)
        for n in (60_000, 900):
            pts = [rng.uniform(0, 4.0, n), rng.uniform(0, 3.0, n), rng.uniform(-0.5, 0.5, n)]
            dev = [DeviceArray(a.astype(np.float32)) for a in pts]
            ref = [DeviceArray(np.zeros(n, np.float32)) for _ in range(3)]
            got = [DeviceArray(np.full(n, np.nan, np.float32)) for _ in range(3)]
            synchronize()
            S.velocity_at_device(n, *[d.ptr for d in dev], *[o.ptr for o in ref])
            S.velocity_at_device_noalloc(n, *[d.ptr for d in dev], *[o.ptr for o in got])
            S.synchronize()
            for a, b in zip(got, ref):
                ha, hb = a.download(), b.download()
                assert np.all(np.isfinite(hb)) and np.array_equal(ha, hb), n
        with pytest.raises(RbcError, match="no-allocation scratch holds"):
            S.velocity_at_device_noalloc(1_398_102, *[d.ptr for d in dev], *[o.ptr for o in got])
