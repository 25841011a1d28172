"""Fluctuation-dissipation pin of the oracle's Brown thermal field
(modules/mod_brown.go:124-140, src/libhotspin/temperature.cu:14-71): an
ensemble of 1024 independent macrospin cells reaches the stationary state
that the reference's SDE predicts in closed form (tests/fdt_util.py):

- LLG limit: <cos theta> = L(Ms0 V B / (kB T (1 + lambda^2))) at 150, 300 and
  600 K; with a non-cubic cell and cutoff_dt = 2 dt (the noise variance then
  scales as dt / cutoff_dt, mod_brown.go:136);
- LLB-r with the longitudinal field: <cos theta> = L(xi(|m|)) at 900 and
  1000 K in a 0.3 T field, and at B = 0 the mean |m| and its variance
  against the stationary radial density (mean = the Brillouin m_e(T) up to
  O(lambda^2));
- a deliberately doubled prefactor, planted in the oracle, fails each check.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import fdt_util as F

N = (1, 32, 32)
NCELL = 1024
CUBE = (2e-9, 2e-9, 2e-9)


@pytest.fixture(autouse=True)
def _threads():
    po.set_threads(8)
    po.set_thermal_prefactor_scale(1.0)
    yield
    po.set_thermal_prefactor_scale(1.0)


def _sample(kw, cell, dt, nequil, nsample, m0):
    o = po.LLB(N, cell, **kw)
    m = np.zeros((3, NCELL))
    m[0] = m0
    def advance(i0, n):
        for i in range(i0, i0 + n):
            o.step("rk4", m, dt, i)

    return F.run(advance, lambda: m, NCELL, nequil, nsample)


def _llg(T, cell=CUBE, cutoff_dt=0.0, dt=1e-13, B=1.3, lam=0.1):
    s = _sample(F.params(T, B, lam, cutoff_dt=cutoff_dt), cell, dt, 2000, 2000, 1.0)
    got, se = s.cos()
    want = F.langevin(F.xi_transverse(T, B, cell, lam, 1.0, cutoff_dt, dt))
    return got, se, want


@pytest.mark.parametrize("T", [150.0, 300.0, 600.0])
def test_llg_limit_langevin(T):
    got, se, want = _llg(T)
    assert abs(got - want) <= 3 * se, (got, want, se)


def test_llg_limit_volume_and_cutoff_dt():
    """Non-cubic 2 x 3 x 1.5 nm cells; cutoff_dt = 2 dt halves the noise variance."""
    got, se, want = _llg(600.0, cell=(2e-9, 3e-9, 1.5e-9), cutoff_dt=2e-13)
    assert abs(got - want) <= 3 * se, (got, want, se)
    # and it is not the answer without the cutoff
    assert abs(got - F.langevin(F.xi_transverse(600.0, 1.3, (2e-9, 3e-9, 1.5e-9), 0.1))) > 10 * se


def test_doubled_prefactor_fails_llg():
    po.set_thermal_prefactor_scale(2.0)
    got, se, want = _llg(300.0)
    assert abs(got - want) > 10 * se, (got, want, se)


def _h_lf(T):
    def f(r):
        kw = dict(F.params(T, 0.0, longfield=1, normalize=0), thermal=0, zeeman=0, torque=0, local00nc=0)
        o = po.LLB((1, 1, r.size), CUBE, **kw)
        m = np.zeros((3, r.size))
        m[0] = r
        o.rhs(m)
        return o.heff()[0]
    return f


def _llb_transverse(T, B=1.3, lam=0.1):
    s = _sample(F.params(T, B, lam, normalize=0, longfield=1), CUBE, 5e-14, 4000, 4000, F.m_e_half(T))
    got, se = s.cos()
    r, _ = s.radius()
    return got, se, F.langevin(F.xi_transverse(T, B, CUBE, lam, r))


@pytest.mark.parametrize("T", [900.0, 1000.0])
def test_llb_transverse_langevin(T):
    got, se, want = _llb_transverse(T, B=0.3)
    assert abs(got - want) <= 3 * se, (got, want, se)


def _llb_radial(T, lam=0.1):
    s = _sample(F.params(T, 0.0, lam, normalize=0, longfield=1), CUBE, 2.5e-14, 4000, 4000, F.m_e_half(T))
    r, se = s.radius()
    return r, se, s.radial_var(), F.radial_prediction(T, lam, CUBE, _h_lf(T))


@pytest.mark.parametrize("T", [900.0, 1000.0])
def test_llb_radial_equilibrium(T):
    r, se, var, (want_r, want_var, me) = _llb_radial(T)
    assert abs(want_r - me) < 3e-4                   # the density's mean is the MFA m_e(T)
    assert abs(r - want_r) <= 3 * se, (r, want_r, se)
    assert abs(var / want_var - 1) <= 0.08, (var, want_var)


def test_doubled_prefactor_fails_llb():
    po.set_thermal_prefactor_scale(2.0)
    got, se, want = _llb_transverse(1000.0)
    assert abs(got - want) > 10 * se, (got, want, se)
    r, se, var, (want_r, want_var, _) = _llb_radial(900.0)
    assert var / want_var > 1.5, (var, want_var)
