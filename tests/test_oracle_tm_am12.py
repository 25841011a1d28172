"""Oracle pins for the coupled AM12 driver of (Te, Tp, m) (CPU only).

The electron-phonon subsystem gamma Te dTe/dt = G (Tp - Te),
Cl dTp/dt = G (Te - Tp) conserves E = gamma Te^2 / 2 + Cl Tp and relaxes to
the temperature where Te = Tp on that energy (an independent LSODA solve of
the same two-variable ODE pins the trajectory); with diffusion and a laser
pulse, AM12 at tight tolerances must agree with the adaptive RK23 of the
same system.

Spin heating (Qs = ws - qs, the laser recipe's molecular-energy-density rate
minus its dissipation rate) is pinned as the power of the local damping term
k_loc = gamma alpha H: with the torque off the LLB right-hand side is k_loc
alone, and Qs = 3 kB Tc n S/(S+1) m.k_loc + mu0 Ms0 H.k_loc.  The Debye
lattice capacity enters dTp/dt as 9 n kB cp(Tp/Td) with cp the integral that
the Cp pins check against quadrature."""
import math

import numpy as np
from scipy import integrate

from hotspin_amd import workloads as W
from oracle import pyoracle as po

N, CELL = (2, 8, 16), (2e-9, 4e-9, 4e-9)
GAM, CL, G = 6e3, 3.8e6, 1.8e18
KB, MU0 = 1.38064881313131313e-23, 4e-7 * math.pi  # the constants of the reference's headers


def _llb(longfield=1):
    kw = W.llbr_params(T=300.0, seed=5, Bext=(0.0, 0.01, 0.02))
    kw["thermal"] = 0
    kw["longfield"] = longfield
    return po.LLB(N, CELL, **kw)


def _m0(T):
    ncell = N[0] * N[1] * N[2]
    return W.random_state(ncell, W.equilibrium_m(T, 1044.0, 0.5), np.random.default_rng(9))


def test_electron_phonon_relaxation_matches_the_ode_and_conserves_energy():
    q = po.tm_params(gammaE=GAM, Cl=CL, G=G, kappa=0.0, Q0=0.0, sigma_r=1e-6, sigma_t=1e-13, t0=0.0, xc=0.0,
                     yc=0.0, maxErrT=1e-3)
    o = _llb(longfield=0)  # slow transverse dynamics: dt set by the electron-phonon relaxation
    ncell = N[0] * N[1] * N[2]
    Te, Tp = np.full(ncell, 900.0), np.full(ncell, 300.0)
    m = np.zeros((3, ncell))
    m[0] = 0.6  # uniform: no exchange torque, slow dynamics
    m[2] = 0.5
    E0 = GAM * Te[0] ** 2 / 2 + CL * Tp[0]
    T_end = 5e-12
    t, dt, steps = 0.0, 1e-15, 0
    while t < T_end * (1 - 1e-12) and steps < 3000:
        dt = min(dt, T_end - t)
        dt, t, st = o.step_tm_am12(q, m, Te, Tp, t, dt, steps, absErr=(1e-3, 1e-3, 1e-4), relErr=(1e-5, 1e-5, 1e-3),
                                   mindt=1e-20, maxdt=5e-14)
        steps += 1
    assert abs(t - T_end) < 1e-20
    # the same two-variable ODE by an independent integrator
    sol = integrate.solve_ivp(lambda _, y: [G * (y[1] - y[0]) / (GAM * y[0]), G * (y[0] - y[1]) / CL],
                              (0.0, T_end), [900.0, 300.0], method="LSODA", rtol=1e-11, atol=1e-9)
    te_ref, tp_ref = sol.y[0, -1], sol.y[1, -1]
    assert te_ref < 700.0  # well into the relaxation (tau = gamma Te / G ~ 3 ps)
    assert np.abs(Te - te_ref).max() < 0.05 and np.abs(Tp - tp_ref).max() < 0.05
    # E = gamma Te^2/2 + Cl Tp is conserved by the ODE; the trapezoidal rule keeps it to the tolerance
    E = GAM * Te ** 2 / 2 + CL * Tp
    assert np.abs(E / E0 - 1).max() < 1e-5
    assert np.all(np.isfinite(m))


def test_am12_agrees_with_rk23_on_the_pumped_film():
    st_ = 100e-15
    q = po.tm_params(gammaE=GAM, Cl=CL, G=G, kappa=90.0,
                     Q0=GAM * (900.0 ** 2 - 300.0 ** 2) / (2 * st_ * math.sqrt(2 * math.pi)), sigma_r=20e-9,
                     sigma_t=st_, t0=150e-15, xc=32e-9, yc=16e-9, maxErrT=1e-4)
    ncell = N[0] * N[1] * N[2]
    T_end = 300e-15
    res = {}
    for method in ("am12", "rk23"):
        o = _llb()
        m, Te, Tp = _m0(300.0), np.full(ncell, 300.0), np.full(ncell, 300.0)
        t, dt, i = 0.0, 1e-16, 0
        while t < T_end * (1 - 1e-12):
            dt = min(dt, T_end - t)
            if method == "am12":
                dt, t, _ = o.step_tm_am12(q, m, Te, Tp, t, dt, i, absErr=(1e-5, 1e-5, 1e-6), relErr=(1e-7, 1e-7, 1e-6),
                                          mindt=1e-20, maxdt=2e-15)
            else:
                dt, t, _ = o.step_tm(q, m, Te, Tp, t, dt, i, maxErr=1e-7, mindt=1e-20, maxdt=2e-15)
            i += 1
        res[method] = (m.copy(), Te.copy(), Tp.copy())
    (ma, Tea, Tpa), (mr, Ter, Tpr) = res["am12"], res["rk23"]
    assert Tea.max() > 500.0  # the pulse has heated the spot
    assert np.abs(Tea - Ter).max() < 0.05
    assert np.abs(Tpa - Tpr).max() < 0.05
    assert np.abs(ma - mr).max() < 2e-4


def test_spin_heat_is_the_power_of_the_local_damping():
    kw = W.llbr_params(T=300.0, seed=5, Bext=(0.0, 0.01, 0.02))
    kw.update(torque=0)  # k = gamma alpha H only
    o = po.LLB(N, CELL, **kw)
    m, k = o.rhs(_m0(300.0), step=3, attempt=0, dt=1e-15)
    H = o.heff()
    Qs = o.spin_heat(m)
    S = kw["JMul"]
    ws = 3 * KB * kw["TcMul"] * kw["nMul"] * S / (S + 1) * np.sum(m * k, axis=0)
    want = ws + MU0 * kw["msat0T0Mul"] * np.sum(H * k, axis=0)
    assert np.allclose(Qs, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # both parts matter here: the damping power and the longitudinal relaxation term
    assert np.abs(ws).max() > 1e-3 * np.abs(want).max()
    # without the local damping term there is no spin heating
    kw.update(local00nc=0)
    o2 = po.LLB(N, CELL, **kw)
    m2, _ = o2.rhs(_m0(300.0), step=3, attempt=0, dt=1e-15)
    assert np.all(o2.spin_heat(m2) == 0.0)


def test_tm_rhs_debye_capacity_and_spin_heat_source():
    Td, nD = 470.0, 9.14e28
    kw = dict(gammaE=GAM, Cl=CL, G=G, kappa=90.0, Q0=1e21, sigma_r=20e-9, sigma_t=1e-13, t0=0.0, xc=32e-9,
              yc=16e-9, maxErrT=1e-3)
    rng = np.random.default_rng(4)
    ncell = N[0] * N[1] * N[2]
    Te, Tp = 300.0 + 500 * rng.random(ncell), 250.0 + 300 * rng.random(ncell)
    Qs = 1e21 * rng.standard_normal(ncell)
    dte0, dtp0 = po.tm_rhs(po.tm_params(**kw), N, CELL, (0, 0, 0), Te, Tp, 2e-14)
    dte, dtp = po.tm_rhs(po.tm_params(Td=Td, nD=nD, **kw), N, CELL, (0, 0, 0), Te, Tp, 2e-14, Qs)
    cp = po.cp(Tp, None, None, Td)
    assert np.allclose(dtp, G * (Te - Tp) / (9 * nD * KB * cp), rtol=1e-13)
    assert np.allclose(dte, dte0 + Qs / (GAM * Te), rtol=1e-12, atol=1e-12 * np.abs(dte).max())
    # at Dulong-Petit (T >> Td) the Debye capacity approaches 3 n kB
    hot = np.full(ncell, 1e5)
    _, dtp_hot = po.tm_rhs(po.tm_params(Td=Td, nD=nD, **kw), N, CELL, (0, 0, 0), hot + 10.0, hot, 0.0)
    assert np.allclose(dtp_hot, G * 10.0 / (3 * nD * KB), rtol=1e-3)
