"""The laser, ensemble and serving workloads at their own sizes and
parameters, GPU against the FP64 oracle.

- Laser film (256x256x8 cells of 4x4x2 nm, LLB-r + two-temperature model,
  Gaussian pump): the Te/Tp right-hand side at 1e-12 and the LLB right-hand
  side at T = Te at 2e-5 on a heated, non-uniform temperature field; five
  coupled RK23 steps from the workload's initial state replayed from the
  oracle's dt (same accept/reject decisions, max|dm| <= 1e-5, Te and Tp to
  1e-6 K); the pump drives the spot centre's Te to 1.3 Tc (within 1 %).
- Ensemble (32^3, 64x64x4, 128x128x2 and 96x48x8 meshes of 3 nm cells,
  seeds and temperatures of the 512-solve draw, J = 7/2): three solves of
  each shape stepped together by the ensemble graph, 20 RK4 steps at
  dt = 1e-15 s; the middle solve of each shape against the oracle,
  max|dm| <= 1e-5.
- Serving: one request of every shape class of the stream (16^3 up to
  256x256x4) among others through the continuous-batching loop; each result
  equals the request solved alone bit for bit, and each class's request,
  ten adaptive RK23 steps, equals the oracle's own ten steps (the same
  rejections, the next dt proposal to 1e-2, max|dm| <= 1e-5).
The oracle takes about a minute on 16 host threads for the whole file."""
import os

import numpy as np
import pytest

from hotspin_amd import workloads as W
from hotspin_amd._lib import check
from hotspin_amd.api import SLOT_DONE, Batch, Ensemble, SlotBatch, TMSolver
from hotspin_amd.serve import Server
from oracle import pyoracle as po
from tests.gpu_util import relerr, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    po.set_threads(min(16, os.cpu_count() or 1))
    yield
    check("configs")


def test_laser_film_rhs_and_coupled_steps():
    w, kw, tm, m0 = W.laser()
    N, cell = w["N"], w["cell"]
    ncell = N[0] * N[1] * N[2]
    q = po.tm_params(**tm)
    g = TMSolver(N, cell, tm=tm, T0=w["T0"], **kw)
    # right-hand sides on a heated, non-uniform temperature field
    rng = np.random.default_rng(11)
    Te = w["T0"] + 900 * rng.random(ncell)
    Tp = w["T0"] + 200 * rng.random(ncell)
    g.set_m(m0)
    g.set_T(Te, Tp)
    print("configs: laser film %s RHS" % (N,), flush=True)
    for t in (0.0, tm["t0"]):
        dte, dtp = g.tm_rhs(t)
        ote, otp = po.tm_rhs(q, N, cell, (0, 0, 0), Te, Tp, t)
        assert relerr(dte, ote) <= 1e-12
        assert relerr(dtp, otp) <= 1e-12
    kwo = dict(kw)
    kwo.update(TMul=1.0, T=Te)
    o = po.LLB(N, cell, **kwo)
    m32 = m0.astype(np.float32).astype(np.float64)  # the m the GPU's stage reads
    _, ko = o.rhs(m32, step=4, attempt=0, dt=1e-15)
    kg, Hg = g.rhs(step=4, attempt=0, dt=1e-15)
    assert relerr(Hg, o.heff()) <= 2e-5
    assert relerr(kg, ko) <= 2e-5
    del o, kg, Hg, ko
    # coupled RK23 steps from the workload's initial state, replayed from the oracle's dt
    o = po.LLB(N, cell, **kw)
    mo, t, dt = m0.copy(), 0.0, 1e-15
    Te, Tp = np.full(ncell, w["T0"]), np.full(ncell, w["T0"])
    for i in range(5):
        m_b, Te_b, Tp_b, t_b, dt_in = mo.copy(), Te.copy(), Tp.copy(), t, dt
        dt, t, st = o.step_tm(q, mo, Te, Tp, t, dt, i, maxErr=w["tol"], mindt=1e-20, maxdt=5e-14)
        g.set_m(m_b)
        g.set_T(Te_b, Tp_b)
        g.time = t_b
        print("configs: laser step %d dt %.3e" % (i, dt_in), flush=True)
        dtg, stg = g.run("rk23", 1, dt_in, maxErr=w["tol"], mindt=1e-20, maxdt=5e-14, step0=i)
        assert stg.rejected == int(st[3]) - 1, (i, stg.rejected, st)
        assert abs(dtg / dt - 1) <= 1e-3, (i, dtg, dt)
        assert abs(g.time - t) <= 1e-9 * t
        teg, tpg = g.get_T()
        assert np.abs(g.get_m() - mo).max() <= 1e-5, i
        assert np.abs(teg - Te).max() <= 1e-6, (i, np.abs(teg - Te).max())
        assert np.abs(tpg - Tp).max() <= 1e-6, i
    g.close()


def test_laser_pump_reaches_its_target_peak():
    """Q0 is calibrated on the spot centre's two-temperature ODE; the film's
    own peak Te (read every 2 steps through the pulse) lands on 1.3 Tc."""
    w, kw, tm, m0 = W.laser()
    ncell = w["N"][0] * w["N"][1] * w["N"][2]
    g = TMSolver(w["N"], w["cell"], tm=tm, T0=w["T0"], **kw)
    g.set_m(m0)
    dt, step, peak = 1e-15, 0, 0.0
    print("configs: laser pump to t = %.2g s" % (tm["t0"] + 5 * tm["sigma_t"]), flush=True)
    while g.time < tm["t0"] + 5 * tm["sigma_t"]:
        dt, st = g.run("rk23", 2, dt, maxErr=w["tol"], mindt=1e-20, maxdt=5e-14, step0=step)
        step += st.steps
        peak = max(peak, float(g.get_T()[0].max()))
        assert step < 4000
    g.close()
    assert abs(peak / w["Tmax"] - 1) <= 0.01, (peak, w["Tmax"])
    assert ncell == 524288


def test_ensemble_shapes_from_inside_the_mix_match_oracle():
    mat = W.MATERIAL_WIDE_T
    cfg = W.ensemble(512, seed=0)
    dt, nsteps = 1e-15, 20
    batches = []
    for shape in W.ENSEMBLE_SHAPES:
        qs = [c for c in cfg if c["N"] == shape][:3]
        b = Batch(shape, W.ENSEMBLE_CELL, len(qs), TMuls=[c["T"] for c in qs], seeds=[c["seed"] for c in qs],
                  **W.llbr_params(T=300.0, seed=0, material=mat))
        m0 = np.stack([W.random_domains(shape, 8, W.equilibrium_m(c["T"], mat["Tc"], mat["J"]),
                                        np.random.default_rng(c["seed"])) for c in qs])
        b.set_m(m0)
        batches.append((shape, qs, b, m0))
    e = Ensemble([b for _, _, b, _ in batches])
    print("configs: ensemble of %d shapes, %d RK4 steps" % (len(batches), nsteps), flush=True)
    e.run_graph(nsteps, dt, step0=0)
    for shape, qs, b, m0 in batches:
        got = b.get_m()
        assert np.isfinite(got).all()
        c = qs[1]
        o = po.LLB(shape, W.ENSEMBLE_CELL, **W.llbr_params(T=c["T"], seed=c["seed"], material=mat))
        mo = m0[1].copy()
        for i in range(nsteps):
            o.step("rk4", mo, dt, i)
        err = np.abs(got[1] - mo).max()
        print("configs: ensemble %s T=%.0f K seed %d max|dm| %.2e" % (shape, c["T"], c["seed"], err), flush=True)
        assert err <= 1e-5, (shape, err)
    e.close()
    for _, _, b, _ in batches:
        b.close()


def test_serving_every_shape_class_alone_in_mix_and_vs_oracle():
    reqs = W.requests(n=60, rate=2000.0, seed=3)
    # one request of every class with ten steps, the rest a mix of classes and lengths
    rng = np.random.default_rng(3)
    for i, r in enumerate(reqs):
        if i < len(W.SERVE_SHAPES):
            r["cls"], r["steps"] = i, 10
        else:
            r["cls"], r["steps"] = int(rng.integers(0, 4)), int(rng.integers(5, 30))
        r["N"] = W.SERVE_SHAPES[r["cls"]]
    srv = Server(slots_cells=1 << 18, min_slots=2, max_slots=8, chunk=4)
    try:
        print("configs: serving %d requests" % len(reqs), flush=True)
        res = srv.run(reqs)
        summ = srv.summary()
    finally:
        srv.close()
    assert summ["requests"] == len(reqs) and summ["failed"] == 0
    assert all(res[r["id"]]["steps"] == r["steps"] for r in reqs)
    for r in reqs[:len(W.SERVE_SHAPES)]:
        shape = W.SERVE_SHAPES[r["cls"]]
        m0 = srv._initial_state(r)
        b = SlotBatch(shape, W.SERVE_CELL, 1, **srv.kw)
        b.admit(0, m0, r["T"], r["seed"], r["steps"], srv.dt0, maxErr=srv.tol, mindt=srv.mindt, maxdt=srv.maxdt,
                Bext=r["Bext"])
        st = b.run_to_end(chunk=5)[0]
        assert st.state == SLOT_DONE
        got = res[r["id"]]
        mm = b.get_m(0)
        b.close()
        assert (got["steps"], got["rejected"], got["t_sim"]) == (st.accepted, st.rejected, st.t)
        assert np.array_equal(got["m"], mm), r["id"]
        kw = W.llbr_params(T=r["T"], seed=r["seed"], material=srv.material, Bext=r["Bext"])
        o = po.LLB(shape, W.SERVE_CELL, **kw)
        mo, dto, attempts = m0.copy(), srv.dt0, 0
        for i in range(r["steps"]):
            dto, sto = o.step("rk23", mo, dto, i, maxErr=srv.tol, mindt=srv.mindt, maxdt=srv.maxdt)
            attempts += int(sto[1])
        err = np.abs(mm - mo).max()
        print("configs: serving class %d %s T=%.0f K max|dm| %.2e dt %.4e / %.4e" % (
            r["cls"], shape, r["T"], err, st.h, dto), flush=True)
        assert attempts - r["steps"] == st.rejected, (shape, attempts, st.rejected)
        assert abs(st.h / dto - 1) <= 1e-2, (shape, st.h, dto)
        assert err <= 1e-5, (shape, err)
