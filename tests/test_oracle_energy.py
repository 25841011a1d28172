"""Oracle pins for the energy terms (micromag/energy, modules/mod_energy.go:
25-61,98-116), CPU only.

Closed forms fix each term and its weight: the Zeeman energy of a uniform
state is -V Ms N (m . B_ext); a Fourier mode m = a (cos kx, sin kx, 0) along
a periodic axis is an eigenvector of the 6-neighbour Laplacian with
eigenvalue 2 (cos k dx - 1) / dx^2, so E_ex = -(1/2) V mu0 Ms^2 lex^2
2 (cos k dx - 1) / dx^2 a^2 N; a uniformly magnetised cube of cubic cells has
the demag factor 1/3, so E_demag = (1/2) mu0 Ms^2 |m|^2 V N / 3; the
longitudinal field vanishes at |m| = m_e(T), and with it E_lf, which is
positive above m_e and negative below."""
import math

import numpy as np

from hotspin_amd import workloads as W
from oracle import pyoracle as po

MU0 = 4e-7 * math.pi
MS = 8e5


def _solver(N, cell, **kw):
    base = dict(msat0T0Mul=MS, exchange=0, demag=0, zeeman=0, longfield=0, thermal=0, torque=1, local00nc=0,
                gammaLL=2.211e5)
    base.update(kw)
    return po.LLB(N, cell, **base)


def test_zeeman_energy_of_a_uniform_state():
    N, cell = (2, 3, 4), (2e-9, 3e-9, 4e-9)
    B = (0.01, -0.02, 0.05)
    o = _solver(N, cell, zeeman=1, Bext=B)
    u = np.array([0.3, -0.4, 0.5])
    m = np.repeat(u[:, None], 24, axis=1)
    mr, _ = o.rhs(m)
    E = o.energy(mr)
    V = cell[0] * cell[1] * cell[2]
    assert math.isclose(E[0], -V * MS * 24 * float(u @ np.array(B)), rel_tol=1e-13)
    assert E[1] == E[2] == E[3] == 0.0 and E[4] == E[0]


def test_exchange_energy_of_a_fourier_mode():
    n, dx = 64, 3e-9
    lex = 5e-9
    o = _solver((1, 1, n), (2e-9, 2e-9, dx), pbc=(0, 0, 1), exchange=1, lexMul=lex)
    kq = 2 * math.pi * 5 / (n * dx)
    x = np.arange(n) * dx
    a = 0.8
    m = np.stack([a * np.cos(kq * x), a * np.sin(kq * x), np.zeros(n)])
    mr, _ = o.rhs(m)
    E = o.energy(mr)
    V = 2e-9 * 2e-9 * dx
    lam = 2 * (math.cos(kq * dx) - 1) / dx ** 2
    want = -0.5 * V * MU0 * MS * MS * lex ** 2 * lam * a * a * n
    assert math.isclose(E[1], want, rel_tol=1e-12)
    assert E[1] > 0  # a twisted state costs exchange energy


def test_demag_energy_of_a_uniform_cube():
    n, c = 8, 2e-9
    o = _solver((n, n, n), (c, c, c), demag=1)
    u = np.array([0.0, 0.6, 0.8])
    m = np.repeat(u[:, None], n ** 3, axis=1)
    mr, _ = o.rhs(m)
    E = o.energy(mr)
    want = 0.5 * MU0 * MS ** 2 * float(u @ u) * (c ** 3 * n ** 3) / 3
    assert abs(E[2] / want - 1) < 1e-3, E[2] / want  # Kernel_Arne at accuracy 6: -4.6e-4


def test_longitudinal_energy_vanishes_at_equilibrium():
    T, Tc, S = 300.0, 1044.0, 0.5
    kw = W.llbr_params(T=T, seed=1, Bext=(0.0, 0.0, 0.0))
    kw.update(exchange=0, demag=0, zeeman=0, thermal=0)
    N = (1, 2, 4)
    me = W.equilibrium_m(T, Tc, S)
    u = np.array([0.6, 0.0, 0.8])
    res = {}
    for f in (0.95, 1.0, 1.05):
        o = po.LLB(N, (2e-9,) * 3, **kw)
        mr, _ = o.rhs(np.repeat((f * me * u)[:, None], 8, axis=1))
        res[f] = o.energy(mr)[3]
    assert abs(res[1.0]) < 1e-6 * abs(res[1.05])
    assert res[1.05] > 0 > res[0.95]
