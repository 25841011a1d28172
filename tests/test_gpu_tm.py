"""Laser-heated LLB-r coupled to the two-temperature model, GPU vs the FP64
oracle: the Te/Tp right-hand side (FP64 on both sides, 1e-12 relative), the
LLB right-hand side with T = Te (FP32 fields, 2e-5), and adaptive RK23 steps
of (m, Te, Tp) replayed from the oracle's dt sequence (same accept/reject
decisions, max|dm| <= 1e-5, |dTe|, |dTp| <= 1e-6 K).

With the Debye lattice capacity and spin heating on, the spin term of Qe
comes from the FP32 field of the LLB stages: dTe/dt is held to 2e-5 of
max|Qs|/(gamma Te) (the FP32 field bar) and a step's Te, Tp to 2e-5 of the
step's largest Te change."""
import math

import numpy as np
import pytest

from hotspin_amd import workloads as W
from hotspin_amd.api import TMSolver
from hotspin_amd._lib import check
from oracle import pyoracle as po
from tests.gpu_util import relerr, require_gpu

pytestmark = pytest.mark.gpu

N, CELL = (2, 16, 32), (2e-9, 4e-9, 4e-9)


EXTRA = dict(Td=470.0, nD=9.14e28, spin_heat=1)  # Debye lattice capacity + spin heating


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("tm")


def _problem():
    gam, T0, Tmax, st = 6e3, 300.0, 900.0, 100e-15
    tm = dict(gammaE=gam, Cl=3.8e6, G=1.8e18, kappa=90.0,
              Q0=gam * (Tmax ** 2 - T0 ** 2) / (2 * st * math.sqrt(2 * math.pi)), sigma_r=30e-9, sigma_t=st,
              t0=200e-15, xc=64e-9, yc=32e-9, maxErrT=1e-3)
    kw = W.llbr_params(T=T0, seed=5, Bext=(0.0, 0.01, 0.02))
    rng = np.random.default_rng(9)
    ncell = N[0] * N[1] * N[2]
    m = W.random_state(ncell, W.equilibrium_m(T0, 1044.0, 0.5), rng)
    Te = T0 + 400 * rng.random(ncell)
    Tp = T0 + 100 * rng.random(ncell)
    return tm, kw, m, Te, Tp


@pytest.mark.parametrize("pbc", [(0, 0, 0), (0, 1, 1)])
def test_tm_and_llb_rhs_match_oracle(pbc):
    tm, kw, m, Te, Tp = _problem()
    g = TMSolver(N, CELL, pbc, tm=tm, T0=300.0, **kw)
    g.set_m(m)
    g.set_T(Te, Tp)
    q = po.tm_params(**tm)
    for t in (0.0, 180e-15, 400e-15):
        dte, dtp = g.tm_rhs(t)
        ote, otp = po.tm_rhs(q, N, CELL, pbc, Te, Tp, t)
        assert relerr(dte, ote) <= 1e-12
        assert relerr(dtp, otp) <= 1e-12
    kwo = dict(kw)
    kwo.update(TMul=1.0, T=Te)
    o = po.LLB(N, CELL, pbc, **kwo)
    _, ko = o.rhs(m, step=4, attempt=0, dt=1e-15)
    kg, Hg = g.rhs(step=4, attempt=0, dt=1e-15)
    assert relerr(Hg, o.heff()) <= 2e-5
    assert relerr(kg, ko) <= 2e-5


@pytest.mark.parametrize("extra", [{}, EXTRA], ids=["cl", "debye_spinheat"])
def test_coupled_rk23_replays_oracle_steps(extra):
    tm, kw, m, _, _ = _problem()
    tm.update(extra)
    T0 = 300.0
    ncell = N[0] * N[1] * N[2]
    Te, Tp = np.full(ncell, T0), np.full(ncell, T0)
    q = po.tm_params(**tm)
    o = po.LLB(N, CELL, **kw)
    g = TMSolver(N, CELL, tm=tm, T0=T0, **kw)
    mo, t, dt = m.copy(), 0.0, 1e-15
    for i in range(30):
        m_b, Te_b, Tp_b, t_b, dt_in = mo.copy(), Te.copy(), Tp.copy(), t, dt
        dt, t, st = o.step_tm(q, mo, Te, Tp, t, dt, i, maxErr=1e-5, mindt=1e-20, maxdt=5e-14)
        g.set_m(m_b)
        g.set_T(Te_b, Tp_b)
        g.time = t_b
        dtg, stg = g.run("rk23", 1, dt_in, maxErr=1e-5, mindt=1e-20, maxdt=5e-14, step0=i)
        assert stg.rejected == int(st[3]) - 1, (i, stg.rejected, st)
        assert abs(dtg / dt - 1) <= 1e-2, (i, dtg, dt)  # observed up to 6.6e-3 (spin heating on)
        assert abs(g.time - t) <= 1e-9 * t
        teg, tpg = g.get_T()
        tb = 1e-6 + (2e-5 * np.abs(Te - Te_b).max() if extra else 0.0)
        assert np.abs(g.get_m() - mo).max() <= 1e-5, i
        assert np.abs(teg - Te).max() <= tb, (i, np.abs(teg - Te).max())
        assert np.abs(tpg - Tp).max() <= tb, i
    # the pump has heated the spot (less with spin heating: demagnetising takes energy from the electrons)
    assert Te.max() > (400.0 if extra else 450.0)


@pytest.mark.parametrize("extra", [{}, EXTRA], ids=["cl", "debye_spinheat"])
def test_coupled_am12_matches_oracle_step_by_step(extra):
    """AM12 of (Te, Tp, m) (the reference's laser recipe solver) from the same
    state and dt each step: the same fixed-point iteration counts, accepted dt
    within 1e-3, max|dm| <= 1e-5 and |dTe|, |dTp| <= 1e-6 K."""
    tm, kw, m, _, _ = _problem()
    tm.update(extra)
    T0 = 300.0
    ncell = N[0] * N[1] * N[2]
    Te, Tp = np.full(ncell, T0), np.full(ncell, T0)
    q = po.tm_params(**tm)
    o = po.LLB(N, CELL, **kw)
    g = TMSolver(N, CELL, tm=tm, T0=T0, **kw)
    ab, rl = (1e-3, 1e-3, 1e-4), (1e-5, 1e-5, 1e-3)
    g.set_implicit(maxAbsErr=ab[2], maxRelErr=rl[2])
    g.set_implicit_T(maxAbsErr=ab[0], maxRelErr=rl[0])
    mo, t, dt = m.copy(), 150e-15, 2e-16
    for i in range(12):
        m_b, Te_b, Tp_b, t_b, dt_in = mo.copy(), Te.copy(), Tp.copy(), t, dt
        dt, t, st = o.step_tm_am12(q, mo, Te, Tp, t, dt, i, absErr=ab, relErr=rl, mindt=1e-20, maxdt=5e-14)
        g.set_m(m_b)
        g.set_T(Te_b, Tp_b)
        g.time = t_b
        dtg, stg = g.run("am12", 1, dt_in, mindt=1e-20, maxdt=5e-14, step0=i)
        assert stg.iterations == int(st[4]), (i, stg.iterations, st)
        assert abs(stg.t - st[5]) <= 1e-3 * st[5], (i, stg.t, st)
        # the proposal is dt (0.8/err)^(1/2).  With spin heating this run alternates between
        # steps cut by iterator retries and steps whose Te error estimate (corrector minus
        # predictor) is ~1e-6 of the tolerance: a difference of FP32-field Qs values at that
        # level, so err, and the proposal, are good to ~10% there (perturbing the oracle's
        # Qs by 1e-6 relative moves those proposals by 1-4%)
        floor = extra and min(st[:3]) < 1e-4
        assert abs(dtg - dt) <= (0.1 if floor else 1e-3) * dt, (i, dtg, dt, st)  # observed 7.1e-2 / 7.8e-4
        teg, tpg = g.get_T()
        tb = 1e-6 + (2e-5 * np.abs(Te - Te_b).max() if extra else 0.0)
        assert np.abs(g.get_m() - mo).max() <= 1e-5, i
        assert np.abs(teg - Te).max() <= tb, (i, np.abs(teg - Te).max())
        assert np.abs(tpg - Tp).max() <= tb, i
        # continue both from the oracle's state
        Te, Tp = Te.copy(), Tp.copy()
    assert Te.max() > 300.2  # the pump is on


def test_tm_rhs_with_debye_capacity_and_spin_heat():
    tm, kw, m, Te, Tp = _problem()
    tm.update(EXTRA)
    g = TMSolver(N, CELL, tm=tm, T0=300.0, **kw)
    g.set_m(m)
    g.set_T(Te, Tp)
    q = po.tm_params(**tm)
    kwo = dict(kw)
    kwo.update(TMul=1.0, T=Te)
    o = po.LLB(N, CELL, **kwo)
    mo, _ = o.rhs(m, step=0, attempt=0, dt=1e-15)  # the state's RHS that the GPU entry evaluates
    Qs = o.spin_heat(mo)
    gam = tm["gammaE"]
    for t in (0.0, 200e-15):
        dte, dtp = g.tm_rhs(t)
        ote, otp = po.tm_rhs(q, N, CELL, (0, 0, 0), Te, Tp, t, Qs)
        assert relerr(dtp, otp) <= 1e-12  # the Debye integral is FP64 on both sides
        # FP32-field rounding scales with |H|, so with the field's largest Qs (ws and qs
        # cancel in some cells)
        bar = 2e-5 * np.abs(Qs).max() / (gam * Te) + 1e-12 * np.abs(ote).max()
        assert np.all(np.abs(dte - ote) <= bar), np.max(np.abs(dte - ote) / bar)
        # the spin term is well above that bar: leaving it out would fail
        assert np.abs(Qs / (gam * Te)).max() > 1e3 * bar.max()


def test_energy_of_a_two_temperature_state():
    """hs_solver_energy with T = Te per cell (the stages' FP32 copy of Te): the
    oracle evaluated with T = Te rounded to FP32, terms to the bars of
    tests/test_gpu_energy.py."""
    tm, kw, m, Te, Tp = _problem()
    g = TMSolver(N, CELL, tm=tm, T0=300.0, **kw)
    g.set_m(m)
    g.set_T(Te, Tp)
    kwo = dict(kw)
    kwo.update(TMul=1.0, T=Te.astype(np.float32).astype(np.float64))
    o = po.LLB(N, CELL, **kwo)
    mo, _ = o.rhs(m)
    Eo = o.energy(mo)
    Eg = g.energy()
    for t, tol in ((0, 1e-12), (1, 1e-10), (2, 1e-5), (3, 1e-9)):
        assert abs(Eg[t] - Eo[t]) <= tol * abs(Eo[t]), (t, Eg, Eo)
