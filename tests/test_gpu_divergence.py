"""A diverged state is never accepted on the GPU (the oracle's rule,
tests/test_oracle_divergence.py): the stage kernels map a NaN error term to
+inf before the max, a non-finite estimate is retried at factor 0.1 while
retries and dt remain, and then hs_solver_run returns HS_ECONV and a slot
ends HS_SLOT_FAILED (serving then reports the request as failed)."""
import math

import numpy as np
import pytest

from hotspin_amd._lib import HotspinError, check
from hotspin_amd.api import SLOT_DONE, SLOT_FAILED, SlotBatch, Solver
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu
N = (2, 16, 32)
NC = 2 * 16 * 32
CELL = (3e-9, 3e-9, 3e-9)


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("divergence")


def _kw():
    Ms, J = 8e5, 0.5
    return dict(msat0T0Mul=Ms, exchange=1, lexMul=math.sqrt(2 * 1.3e-11 / (4e-7 * math.pi * Ms ** 2)), demag=1,
                zeeman=1, Bext=(0.02, 0.0, 0.01), longfield=1, JMul=J, nMul=Ms / (2 * 9.27400968202020202e-24 * J),
                TcMul=1044.0, TMul=300.0, thermal=1, seed=3, gammaLL=2.211e5, torque=1, local00nc=1,
                alphaMul=(0.02,) * 3)


def _state(seed):
    m = np.random.default_rng(seed).standard_normal((3, NC))
    return m * 0.9 / np.linalg.norm(m, axis=0)


@pytest.mark.parametrize("method", ["rk12", "rk23"])
def test_nan_cell_fails_the_step(method):
    s = Solver(N, CELL, **_kw())
    m = _state(0)
    m[1, 5] = np.nan
    s.set_m(m)
    print("divergence: %s with a NaN cell" % method, flush=True)
    with pytest.raises(HotspinError, match="hs error 10003"):
        s.run(method, 3, 1e-15, maxErr=1e-5, mindt=1e-20, maxdt=1e-12)
    s.close()


@pytest.mark.parametrize("method", ["rk12", "rk23"])
def test_absurd_fixed_dt_ends_in_failure(method):
    s = Solver(N, CELL, **_kw())
    s.set_m(_state(1))
    dt = 1e-9
    print("divergence: %s at dt = mindt = maxdt = %g" % (method, dt), flush=True)
    with pytest.raises(HotspinError, match="hs error 10003"):
        for i in range(50):
            s.run(method, 1, dt, maxErr=1e-5, mindt=dt, maxdt=dt)
            assert np.isfinite(s.get_m()).all(), i  # every accepted step is finite
    s.close()


def test_recoverable_blow_up_is_retried():
    s = Solver(N, CELL, **_kw())
    s.set_m(_state(2))
    print("divergence: rk23 from dt = 1e-6", flush=True)
    dt, st = s.run("rk23", 2, 1e-6, maxErr=1e-5, mindt=1e-20, maxdt=1e-12)
    assert np.isfinite(s.get_m()).all()
    assert st.rejected >= 1 and st.steps == 2
    s.close()


def test_nan_slot_fails_alone():
    b = SlotBatch(N, CELL, 2, **_kw())
    bad = _state(3)
    bad[0, 7] = np.nan
    b.admit(0, bad, 300.0, 3, 5, 1e-15)
    b.admit(1, _state(4), 300.0, 4, 5, 1e-15)
    print("divergence: slot batch with one NaN slot", flush=True)
    st = b.run_to_end()
    assert st[0].state == SLOT_FAILED
    assert st[1].state == SLOT_DONE and st[1].accepted == 5
    assert np.isfinite(b.get_m(1)).all()
    b.close()
