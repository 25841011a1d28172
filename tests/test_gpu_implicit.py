"""Implicit drivers on the GPU (AM12, backward Euler) vs the FP64 oracle.

Each step starts both sides from the same state and dt; the GPU's
fixed-point maxima are FP64 over FP32 slopes, so the iterator decisions
(error <= 0.1 of the tolerance scale, contraction < 1) and the step
controller agree unless a value sits within ~1e-6 relative of a threshold.
Bars: max|dm| <= 1e-5 (normalised for the LLG limit), the same iteration
counts, the accepted and the proposed dt within 1e-3 relative."""
import numpy as np
import pytest

from hotspin_amd.api import Solver
from hotspin_amd._lib import check
from oracle import pyoracle as po
from tests.gpu_util import require_gpu
from tests.test_gpu_solver import llb_problem, llg_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("implicit")


def _norm(m):
    return m / np.linalg.norm(m, axis=0)


def _problem(kind, thermal=False):
    if kind == "llg":
        N, cell = (1, 16, 64), (3e-9, 3.90625e-9, 3.90625e-9)
        kw, m = llg_problem(N, cell)
        return N, cell, kw, m, 1e-13
    N, cell = (2, 16, 32), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, thermal=thermal)
    return N, cell, kw, m, 2e-15


@pytest.mark.parametrize("kind,thermal", [("llg", False), ("llb", False), ("llb", True)])
def test_bdf_euler_matches_oracle(kind, thermal):
    N, cell, kw, m, _ = _problem(kind, thermal)
    dt = 1e-15  # within the fixed-point contraction of both problems (3 updates of 1e-5)
    o = po.LLB(N, cell, **kw)
    g = Solver(N, cell, **kw)
    g.set_m(m)
    mo = m.copy()
    iters = 0
    for i in range(30):
        sto = o.step_bdf_euler(mo, dt, i)
        iters += int(sto[1])
    _, st = g.run("bdf-euler", 30, dt, step0=0)
    mg = g.get_m()
    if kind == "llg":
        mg, mo = _norm(mg), _norm(mo)
    assert np.abs(mg - mo).max() <= 1e-5
    assert st.steps == 30 and st.iterations == iters


def test_bdf_euler_failure_is_an_error_not_a_crash():
    N, cell, kw, m, _ = _problem("llb")
    g = Solver(N, cell, **kw)
    g.set_m(m)
    with pytest.raises(Exception):
        g.run("bdf-euler", 1, 1e-14)  # beyond the fixed-point iteration's contraction limit


@pytest.mark.parametrize("kind,thermal", [("llg", False), ("llb", False), ("llb", True)])
def test_am12_matches_oracle_step_by_step(kind, thermal):
    N, cell, kw, m, dt = _problem(kind, thermal)
    o = po.LLB(N, cell, **kw)
    g = Solver(N, cell, **kw)
    g.set_m(m)
    mo = m.copy()
    for i in range(25):
        dto, sto = o.step_am12(mo, dt, i)
        dtg, st = g.run("am12", 1, dt, step0=i)
        assert st.iterations == int(sto[2]), (i, st.iterations, sto)
        # the same attempts, so the accepted dt differs only where an iterator retry
        # cut it by (0.6/alpha)^(1/2): alpha is a ratio of iterate differences,
        # h |k(y) - k(y')| with FP32 slopes, good to ~1e-4 relative
        assert abs(st.t - sto[4]) <= 1e-3 * sto[4]
        assert abs(dtg - dto) <= 1e-3 * dto
        mg = g.get_m()
        a, b = (_norm(mg), _norm(mo)) if kind == "llg" else (mg, mo)
        assert np.abs(a - b).max() <= 1e-5, i
        # both continue from the oracle's state and dt
        g.set_m(mo)
        dt = dto


def test_implicit_tolerances_are_validated():
    N, cell, kw, m, _ = _problem("llb")
    g = Solver(N, cell, **kw)
    with pytest.raises(Exception):
        g.set_implicit(maxAbsErr=0.0)
    g.set_implicit(1e-5, 1e-4, 1e-6, 5)
