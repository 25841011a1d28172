"""A diverged state is never accepted by the oracle's adaptive drivers.

The error estimates map NaN to +inf (a plain max drops NaN, so a NaN state
would otherwise report a finite, even zero, error and be accepted); a
non-finite estimate is retried at the smallest factor (0.1) while retries
remain and dt can still shrink, and the step fails after that.  The
reference's own reductor (reduce.cu partial maxima, finished on the host)
has no such guard; this is a deliberate divergence from it."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

N = (1, 4, 8)
NC = 32


def _llb(**over):
    Ms, J = 8e5, 0.5
    kw = dict(msat0T0Mul=Ms, exchange=1, lexMul=math.sqrt(2 * 1.3e-11 / (4e-7 * math.pi * Ms ** 2)), demag=1,
              zeeman=1, Bext=(0.02, 0.0, 0.01), longfield=1, JMul=J, nMul=Ms / (2 * 9.27400968202020202e-24 * J),
              TcMul=1044.0, TMul=300.0, thermal=1, seed=3, gammaLL=2.211e5, torque=1, local00nc=1,
              alphaMul=(0.02,) * 3)
    kw.update(over)
    return po.LLB(N, (3e-9, 3e-9, 3e-9), **kw)


def _state(rng):
    m = rng.standard_normal((3, NC))
    return np.ascontiguousarray(m * 0.9 / np.linalg.norm(m, axis=0))


@pytest.mark.parametrize("method", ["rk12", "rk23"])
def test_nan_cell_fails_the_step(method):
    o = _llb()
    m = _state(np.random.default_rng(0))
    m[1, 5] = np.nan
    with pytest.raises(RuntimeError, match="did not converge"):
        o.step(method, m, 1e-15, 0, maxErr=1e-5, mindt=1e-20, maxdt=1e-12)


@pytest.mark.parametrize("method", ["rk12", "rk23"])
def test_absurd_fixed_dt_ends_in_failure_not_a_nan_state(method):
    """dt = 1e-9 s with mindt = maxdt (no room to shrink): the explicit step
    blows up within a few steps, and the step that produces a non-finite
    estimate fails instead of being accepted at mindt."""
    o = _llb()
    m = _state(np.random.default_rng(1))
    dt = 1e-9
    with pytest.raises(RuntimeError, match="did not converge"):
        for i in range(50):
            o.step(method, m, dt, i, maxErr=1e-5, mindt=dt, maxdt=dt)
            assert np.isfinite(m).all(), i  # every accepted step is finite


def test_recoverable_blow_up_is_retried():
    """A too-large first dt whose attempt overflows is retried with smaller
    dt (factor 0.1 per retry) from the finite step-start state."""
    o = _llb()
    m = _state(np.random.default_rng(2))
    dt, st = o.step("rk23", m, 1e-6, 0, maxErr=1e-5, mindt=1e-20, maxdt=1e-12)
    assert np.isfinite(m).all()
    assert st[1] > 1  # at least one rejected attempt
