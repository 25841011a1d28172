"""The continuous-batching request loop: every request completes, and each
result is the same as the request solved alone (bit-exact: per-slot stage
blocks, FFT rows, noise counters and dt control do not depend on what the
other slots hold)."""
import numpy as np
import pytest

from hotspin_amd import workloads as W
from hotspin_amd.api import SLOT_DONE, SlotBatch
from hotspin_amd.serve import Server
from hotspin_amd._lib import check
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

SHAPES = [(2, 16, 32), (4, 8, 16), (1, 32, 32)]


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("serve")


def _requests(n):
    rs = W.requests(n=n, rate=2000.0, seed=7)
    rng = np.random.default_rng(7)
    for r in rs:
        r["cls"] = int(rng.integers(0, len(SHAPES)))
        r["N"] = SHAPES[r["cls"]]
        r["steps"] = int(rng.integers(5, 40))
    return rs


@pytest.mark.parametrize("policy", ["sjf", "fifo"])
def test_serve_results_equal_solves_alone(policy):
    """Both admission policies (least work first under a cap on the running
    cells, or every request at once) complete every request with the result
    of the request solved alone; sjf keeps the running cells within its cap."""
    reqs = _requests(24)
    cap = 3000  # cells: two or three of these requests at a time, the rest queue
    srv = Server(shapes=SHAPES, slots_cells=4096, min_slots=3, max_slots=4, chunk=3, policy=policy,
                 inflight_cells=cap)
    try:
        res = srv.run(reqs)
        summ = srv.summary()
    finally:
        srv.close()
    assert summ["requests"] == 24 and summ["failed"] == 0 and summ["policy"] == policy
    if policy == "sjf":
        assert 0 < srv.peak_running_cells <= cap
    assert all(res[r["id"]]["steps"] == r["steps"] for r in reqs)
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0
    for r in reqs[:: 5]:
        solo = Server(shapes=SHAPES)
        m0 = solo._initial_state(r)
        b = SlotBatch(SHAPES[r["cls"]], W.SERVE_CELL, 1, **srv.kw)
        b.admit(0, m0, r["T"], r["seed"], r["steps"], srv.dt0, maxErr=srv.tol, mindt=srv.mindt, maxdt=srv.maxdt,
                Bext=r["Bext"])
        st = b.run_to_end(chunk=4)[0]
        assert st.state == SLOT_DONE
        got = res[r["id"]]
        assert (got["steps"], got["rejected"]) == (st.accepted, st.rejected)
        assert got["t_sim"] == st.t
        mm = b.get_m(0)
        assert np.isfinite(mm).all()
        assert np.array_equal(got["m"], mm), r["id"]
        b.close()
