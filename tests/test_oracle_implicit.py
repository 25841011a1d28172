"""Oracle pins for the implicit drivers (CPU only): backward Euler (solver/
bdf-euler) and the adaptive AM12 (solver/am12).

A macrospin precessing in a uniform field obeys the linear ODE dm/dt = A m
with A skew-symmetric, so each implicit stage has a closed form: backward
Euler m1 = (I - hA)^-1 m0, the trapezoidal rule (Crank-Nicolson)
m1 = (I - hA/2)^-1 (I + hA/2) m0, a rotation that keeps |m|.  The fixed-point
iterations must land on those within their iteration tolerance, the AM12
error estimate is the corrector-predictor difference, and the step
controller follows dt_new = dt (0.8/err)^(1/2)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

MU0 = 4e-7 * math.pi
GAMMA, BZ = 2.211e5, 0.1
W = GAMMA * BZ / MU0  # precession rate, rad/s


def _macrospin(normalize=0):
    return po.LLB((1, 1, 1), (2e-9, 2e-9, 2e-9), zeeman=1, Bext=(0.0, 0.0, BZ), gammaLL=GAMMA, torque=1,
                  local00nc=0, msat0T0Mul=8e5, normalize=normalize)


def _A():
    """k = gamma m x H with H along internal axis 2: dm0/dt = w m1, dm1/dt = -w m0."""
    return np.array([[0.0, W, 0.0], [-W, 0.0, 0.0], [0.0, 0.0, 0.0]])


def _exact(m0, t):
    c, s = math.cos(W * t), math.sin(W * t)
    return np.array([c * m0[0] + s * m0[1], -s * m0[0] + c * m0[1], m0[2]])


M0 = np.array([0.6, 0.0, 0.8])


def test_rhs_is_the_linear_precession():
    s = _macrospin()
    _, k = s.rhs(M0.reshape(3, 1))
    assert np.allclose(k[:, 0], _A() @ M0, rtol=1e-14, atol=1e-6)


def test_bdf_euler_lands_on_the_implicit_solution():
    s = _macrospin()
    h = 1e-12
    m = np.ascontiguousarray(M0.reshape(3, 1).copy())
    st = s.step_bdf_euler(m, h, 0, maxIterErr=1e-5, maxIter=3)
    want = np.linalg.solve(np.eye(3) - h * _A(), M0)
    # converged to max|y - yb| / max|y + yb| <= 1e-5; the remaining error is the
    # contraction (h w ~ 0.018) times the last update
    assert st[0] <= 1e-5
    assert np.abs(m[:, 0] - want).max() <= 2e-5 * 0.02 * 2
    assert st[1] == 1 + (st[2] - 1)  # one Euler start + one update per further RHS
    # backward Euler shrinks |m| by 1/sqrt(1 + (h w)^2) per step
    assert np.linalg.norm(m[:, 0]) < 1.0


def test_bdf_euler_first_order_convergence():
    s = _macrospin()
    T = 4e-11
    errs = []
    for n in (40, 80):
        m = np.ascontiguousarray(M0.reshape(3, 1).copy())
        for i in range(n):
            s.step_bdf_euler(m, T / n, i, maxIterErr=1e-14, maxIter=60)
        errs.append(np.abs(m[:, 0] - _exact(M0, T)).max())
    assert abs(math.log2(errs[0] / errs[1]) - 1) < 0.15, errs


def test_bdf_euler_iterator_failure_is_reported():
    s = _macrospin()
    m = np.ascontiguousarray(M0.reshape(3, 1).copy())
    with pytest.raises(RuntimeError):  # h w = 1.8: the fixed point iteration diverges
        s.step_bdf_euler(m, 1e-10, 0)


def _cn(h, m0):
    A = _A()
    return np.linalg.solve(np.eye(3) - 0.5 * h * A, (np.eye(3) + 0.5 * h * A) @ m0)


def _be(h, m0):
    return np.linalg.solve(np.eye(3) - h * _A(), m0)


def test_am12_step_is_crank_nicolson_and_controller():
    s = _macrospin()
    absErr, relErr = 1e-4, 1e-3  # the reference's defaults
    m = np.ascontiguousarray(M0.reshape(3, 1).copy())
    dt, st = s.step_am12(m, 2e-13, 0, absErr=absErr, relErr=relErr)
    h = st[4]
    assert h == 2e-13 and st[1] == 1  # first update already within 0.1: no iterator retry
    cn, be = _cn(h, M0), _be(h, M0)
    scale = absErr + relErr * np.abs(be)
    # the iterate stops within 0.1 scale of the previous one and the map contracts
    # by h w ~ 0.004, so it sits within ~1e-3 scale of the fixed point
    assert np.all(np.abs(m[:, 0] - cn) <= 0.01 * scale)
    # step error = corrector - predictor in the same units
    want_err = math.sqrt(np.sum(((cn - be) / scale) ** 2)) / math.sqrt(3)
    assert abs(st[0] - want_err) <= 0.05 * want_err
    assert dt == pytest.approx(h * math.sqrt(0.8 / st[0]), rel=1e-12)
    # the trapezoidal rule is a rotation for the skew A: |m| kept to the iterator tolerance
    assert abs(np.linalg.norm(m[:, 0]) - 1.0) < 1e-5


def test_am12_iterator_retry_cuts_dt():
    """The contraction alpha = tErr / tErr_prev starts from tErr_prev = 0.1, so an
    unconverged first update is an iterator failure (solve_am01.go:113-162): dt is cut
    to dt (0.6/alpha)^(1/2) (or halved) and the step restarts."""
    s = _macrospin()
    m = np.ascontiguousarray(M0.reshape(3, 1).copy())
    dt, st = s.step_am12(m, 1e-10, 0)  # h w = 1.8
    h = st[4]
    assert h < 1e-10 / 2 and st[1] == 1
    assert np.abs(m[:, 0] - _cn(h, M0)).max() <= 0.01 * (1e-4 + 1e-3)
    # the step after an iterator retry is restricted to dt (0.6/alpha)^(1/2) <= the used dt's
    # convergence limit, so it is not larger than 1.5x the used dt here
    assert dt <= 1.5 * h


def test_am12_tracks_the_exact_precession():
    s = _macrospin()
    m = np.ascontiguousarray(M0.reshape(3, 1).copy())
    dt, t, T = 1e-13, 0.0, 5e-11
    while t < T * (1 - 1e-12):
        dt = min(dt, T - t)
        dt, st = s.step_am12(m, dt, 0, absErr=1e-7, relErr=1e-6)
        t += st[4]
    assert np.abs(m[:, 0] - _exact(M0, t)).max() < 2e-5
    assert abs(np.linalg.norm(m[:, 0]) - 1.0) < 1e-6
