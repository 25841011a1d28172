"""Slots (continuous batching) and ensembles of uneven meshes.

- A solve gives the same result alone and in a mix: per-solve stage blocks,
  per-solve FFT rows and per-solve noise counters make it bit-exact, whatever
  the other slots hold and whenever they were admitted or finished.
- Device-side per-slot RK23 control equals the host-driven hs_solver_run
  RK23 of the same solve (same accept/reject sequence; the controller's pow
  runs on the device vs the host, so dt and m agree to 1e-9 relative).
- An ensemble graph step of several batches equals each batch's own graph."""
import numpy as np
import pytest

from hotspin_amd.api import SLOT_DONE, SLOT_IDLE, Batch, Ensemble, SlotBatch, Solver
from hotspin_amd._lib import check
from tests.gpu_util import require_gpu
from hotspin_amd import workloads as W

pytestmark = pytest.mark.gpu

N, CELL = (2, 16, 32), (2e-9, 2e-9, 2e-9)


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("slots")


MAT = W.MATERIAL_WIDE_T


def _kw(thermal=True):
    return W.llbr_params(T=300.0, seed=1234, thermal=thermal, material=MAT, Bext=(0.01, -0.02, 0.05))


def _req(i):
    rng = np.random.default_rng(100 + i)
    ncell = N[0] * N[1] * N[2]
    T = float(rng.uniform(100, 1.2 * MAT["Tc"]))
    m = W.random_state(ncell, max(W.equilibrium_m(T, MAT["Tc"], MAT["J"]), 0.05), rng)
    return dict(m=m, TMul=T, seed=i, Bext=tuple(rng.uniform(-0.1, 0.1, 3)), nsteps=int(rng.integers(5, 25)),
                dt=1e-15)


def _admit(g, slot, r):
    g.admit(slot, r["m"], r["TMul"], r["seed"], r["nsteps"], r["dt"], maxErr=1e-4, mindt=1e-20, maxdt=1e-13,
            Bext=r["Bext"])


def test_solve_alone_equals_solve_in_mix():
    kw = _kw()
    target = _req(0)
    target["nsteps"] = 30
    alone = SlotBatch(N, CELL, 4, **kw)
    _admit(alone, 2, target)
    st = alone.run_to_end(chunk=4)
    ref_m = alone.get_m(2)
    assert np.isfinite(ref_m).all()
    ref = (st[2].accepted, st[2].rejected, st[2].t, st[2].h)
    assert st[2].state == SLOT_DONE and st[2].accepted == 30

    mix = SlotBatch(N, CELL, 4, **kw)
    _admit(mix, 1, _req(1))
    _admit(mix, 3, _req(2))
    mix.advance(3)
    mix.status()
    _admit(mix, 0, target)  # admitted while others are mid-flight
    pending = [_req(i) for i in range(3, 9)]
    while True:
        mix.advance(2)
        st = mix.status()
        for q in (1, 2, 3):
            if st[q].state in (SLOT_DONE, SLOT_IDLE) and pending:
                if st[q].state == SLOT_DONE:
                    mix.get_m(q)
                    mix.release(q)
                _admit(mix, q, pending.pop())
        if st[0].state == SLOT_DONE:
            break
    got = (st[0].accepted, st[0].rejected, st[0].t, st[0].h)
    assert got == ref
    assert np.array_equal(mix.get_m(0), ref_m)


def test_slot_rk23_equals_host_rk23():
    kw = _kw(thermal=False)
    m = W.random_state(N[0] * N[1] * N[2], 0.9, np.random.default_rng(3))
    s = Solver(N, CELL, **kw)
    s.set_m(m)
    dt_h, sth = s.run("rk23", 25, 1e-15, maxErr=1e-5, mindt=1e-20, maxdt=1e-12)
    g = SlotBatch(N, CELL, 2, **kw)
    g.admit(1, m, kw["TMul"], kw["seed"], 25, 1e-15, maxErr=1e-5, mindt=1e-20, maxdt=1e-12, Bext=kw["Bext"])
    st = g.run_to_end(chunk=5)[1]
    assert st.accepted == 25 and st.rejected == sth.rejected
    assert abs(st.h / dt_h - 1) <= 1e-9
    assert abs(st.t / sth.t - 1) <= 1e-9
    assert np.abs(g.get_m(1) - s.get_m()).max() <= 1e-9


def test_ensemble_graph_equals_batch_graphs():
    shapes = [(2, 16, 32), (4, 8, 16)]
    runs = []
    for use_ens in (False, True):
        bs = []
        for i, sh in enumerate(shapes):
            kw = _kw()
            rng = np.random.default_rng(5 + i)
            ncell = sh[0] * sh[1] * sh[2]
            T = rng.uniform(100, 1000, 3)
            m = np.stack([W.random_state(ncell, W.equilibrium_m(t, MAT["Tc"], MAT["J"]), rng) for t in T])
            b = Batch(sh, CELL, 3, TMuls=T, seeds=np.arange(3, dtype=np.uint64) + 3 * i, **kw)
            b.set_m(m)
            bs.append(b)
        if use_ens:
            e = Ensemble(bs)
            e.run_graph(6, 2e-15)
        else:
            for b in bs:
                b.run_graph(6, 2e-15)
        runs.append([b.get_m() for b in bs])
    for a, b in zip(*runs):
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


def test_energy_refuses_a_handle_in_slot_mode():
    """The state of a slot batch lives in its slot buffers, not in the whole-batch
    state that hs_solver_energy reads: the call is an error (HS_EINVAL), not a
    read of stale memory."""
    from hotspin_amd._lib import lib
    g = SlotBatch(N, CELL, 2, **_kw(thermal=False))
    _admit(g, 0, _req(0))
    E = np.zeros(10)
    rc = lib().hs_solver_energy(g._h, E.ctypes.data)
    assert rc == 10001  # HS_EINVAL
    lib().hs_clear_error()
