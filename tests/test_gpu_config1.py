"""The 128x32x1 LLG-limit film at full length: 10,000 fixed RK4 steps of
dt = 1e-13 s (1 ns, a field-driven reversal) on the GPU (FP32 fields, FP64
state, hipGraph replay) against the FP64 oracle's run of the same problem
(tests/golden/config1_rk4.npz, written by tests/golden/make_config1.py).
Bar: max|dm| <= 1e-5 on the final state and on <m>(t) every 100 steps."""
import os

import numpy as np
import pytest

from hotspin_amd import workloads as W
from hotspin_amd.api import Solver
from hotspin_amd._lib import check
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(__file__), "golden", "config1_rk4.npz")


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("config1")


def test_thin_film_10k_rk4_steps_match_oracle():
    ref = np.load(FIX)
    N, cell, kw, m0 = W.thin_film(0)
    s = Solver(N, cell, **kw)
    s.set_m(m0)
    every, dt = int(ref["every"]), float(ref["dt"])
    trace = [m0.mean(axis=1)]
    for _ in range(int(ref["steps"]) // every):
        s.run_graph(every, dt)
        m = s.get_m()
        trace.append((m / np.linalg.norm(m, axis=0)).mean(axis=1))
    m = s.get_m()
    d_final = np.abs(m - ref["m_final"]).max()
    d_trace = np.abs(np.array(trace) - ref["avg_trace"]).max()
    print("config1: max|dm| final %.3e, <m>(t) %.3e" % (d_final, d_trace))
    assert np.isfinite(m).all()
    assert d_trace <= 1e-5
    assert d_final <= 1e-5
