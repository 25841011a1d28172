"""Oracle pins for the MFA helpers, uniaxial anisotropy, nonlocal damping,
divMulPow and wavg (CPU only).

Each check comes from the mathematics rather than from the oracle's own
formula: the inverse of the Brillouin root (Ts), a numerical derivative of the
mean-field self-consistency (kappa), the Debye integral by adaptive
quadrature and its two limits (Cp), the gradient of the anisotropy energy,
the Laplacian's Fourier eigenvalue (nonlocal damping), and the special
cases of the weighted average."""
import math

import numpy as np
import pytest
from scipy import integrate, optimize

from oracle import pyoracle as po

TC = 1044.0


def _bs(S, x):
    """Brillouin function B_S(x), written independently (numpy)."""
    a, b = (2 * S + 1) / (2 * S), 1 / (2 * S)
    return a / np.tanh(a * x) - b / np.tanh(b * x)


def _me(T, S):
    out = np.empty(len(T))
    Tarr = np.ascontiguousarray(T, dtype=np.float64)
    po.lib().or_brillouin(out.ctypes.data, None, Tarr.ctypes.data, None, None, 1.0, 1.0, TC, S, len(T))
    return out


@pytest.mark.parametrize("S", [0.5, 1.0, 3.5])
def test_ts_inverts_the_equilibrium_magnetization(S):
    """Ts(m_e(T)) = T: the spin temperature of the equilibrium |m| is T."""
    # above ~100 K: at lower T, m_e rounds to exactly 1 and Ts takes the |m| = msat0T0 branch (0 K)
    T = np.array([150.0, 200.0, 500.0, 800.0, 1000.0, 1040.0])
    me = _me(T, S)
    Ts = po.ts(len(T), me, None, None, None, 1.0, 1.0, TC, S)
    assert np.max(np.abs(Ts - T) / T) < 1e-6
    # special cases: |m| = msat0T0 -> 0 K, |m| = 0 -> Tc, no material -> 0
    sp = po.ts(3, np.array([1.0, 0.0, 0.5]), np.array([1.0, 1.0, 0.0]), None, None, 1.0, 1.0, TC, S)
    assert sp[0] == 0.0 and sp[1] == TC and sp[2] == 0.0


@pytest.mark.parametrize("S", [0.5, 3.5])
def test_kappa_is_the_mean_field_susceptibility(S):
    """kappa * preS*3Tc*n / u^2 = dm/dh of m = B_S(b (m + h)) at h = 0, b = preS*3Tc/T."""
    preS = S / (S + 1)
    T = np.array([200.0, 600.0, 950.0])
    me = _me(T, S)
    nval = 1.7
    k = po.kappa(T, None, None, np.full(3, nval), me, None, 1.0, 1.0, TC, S)
    for Ti, mi, ki in zip(T, me, k):
        b = preS * 3 * TC / Ti

        def m_of(h):
            return optimize.brentq(lambda m: m - _bs(S, b * (m + h)), 0.02, 1.0, xtol=1e-16, rtol=1e-15)

        eps = 1e-4 * mi
        dmdh = (m_of(eps) - m_of(-eps)) / (2 * eps)
        assert abs(ki * preS * 3 * TC * nval - dmdh) <= 1e-5 * dmdh
    assert po.kappa(np.array([0.0]), None, None, None, np.array([0.5]), None, 1.0, 1.0, TC, S)[0] == 0.0


def test_cp_debye_integral_and_limits():
    def exact(t):  # (T/Td)^3 * int_0^{Td/T} x^4 e^x/(e^x-1)^2 dx
        x = 1.0 / t
        v, _ = integrate.quad(lambda y: y ** 4 * math.exp(y) / math.expm1(y) ** 2 if y > 0 else 0.0, 0, x,
                              epsabs=0, epsrel=1e-12, limit=200)
        return v * t ** 3

    t = np.array([0.2, 0.5, 1.0, 3.0])
    cp = po.cp(t * 470.0, None, None, 470.0)
    for ti, ci in zip(t, cp):
        # 61-interval trapezoid rule: error ~ h^2/12 * f''
        assert abs(ci - exact(ti)) <= 2e-3 * exact(ti)
    # Dulong-Petit: the integrand -> x^2, integral X^3/3, so Cp -> 1/3 (the caller's 9 n kB
    # makes it 3 n kB); the 61-interval trapezoid rule over x^2 overshoots by 1/(2*61^2)
    assert abs(po.cp(np.array([1e6]), None, None, 470.0)[0] - (1 + 1 / (2 * 61 ** 2)) / 3) < 1e-8
    # below T/Td = 0.01: the T^3 law with the integral to infinity, 4 pi^4 / 15
    lo = po.cp(np.array([0.005 * 470.0]), None, None, 470.0)[0]
    assert abs(lo - 4 * math.pi ** 4 / 15 * 0.005 ** 3) < 1e-12 * lo
    # monotone in T, 0 without a Debye temperature, scales with n
    c = po.cp(np.linspace(5.0, 2000.0, 400), None, None, 470.0)
    assert np.all(np.diff(c) > 0)
    assert po.cp(np.array([300.0]), np.array([0.0]), None, 470.0)[0] == 0.0
    assert po.cp(np.array([300.0]), None, np.array([2.5]), 470.0)[0] == pytest.approx(2.5 * exact(300 / 470.0), 2e-3)


def test_energy_flow_linear_and_orthogonal():
    rng = np.random.default_rng(3)
    m = rng.standard_normal((3, 50))
    R = rng.standard_normal((3, 50))
    w = po.energy_flow(m, R, None, None, None, 0.5)
    # linear in R; zero for R perpendicular to m
    assert np.allclose(po.energy_flow(m, 3 * R, None, None, None, 0.5), 3 * w, rtol=1e-14, atol=0)
    Rp = np.cross(m.T, rng.standard_normal((50, 3))).T
    assert np.abs(po.energy_flow(m, Rp, None, None, None, 0.5)).max() < 1e-14
    # S = 1/2, n = Tc = 1: S^2/(S(S+1)) = 1/3 of m.R
    assert np.allclose(w, (m * R).sum(0) / 3, rtol=1e-14, atol=1e-15)
    z = po.energy_flow(m, R, None, None, np.zeros(50), 0.5)
    assert np.all(z == 0.0)


def test_uniaxial_anisotropy_is_the_energy_gradient():
    """For a unit axis, H = -(1/(mu0 Ms)) dE/dm with E = -Ku (m.u)^2."""
    mu0 = 4e-7 * math.pi
    Ku, Ms = 5e5, 1.7e6
    rng = np.random.default_rng(4)
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    m = rng.standard_normal((3, 20))
    h = po.uniaxial_anisotropy(m, None, np.full(20, Ms), 2 * Ku / mu0, [None] * 3, tuple(u))
    for c in range(20):
        g = np.empty(3)
        for a in range(3):
            d = np.zeros(3)
            d[a] = 1e-6
            e = lambda mm: -Ku * (mm @ u) ** 2  # noqa: E731
            g[a] = (e(m[:, c] + d) - e(m[:, c] - d)) / 2e-6
        assert np.allclose(h[:, c], -g / (mu0 * Ms), rtol=1e-7, atol=0)
    # the projection uses the normalized axis, the direction the axis as given
    h2 = po.uniaxial_anisotropy(m, None, np.full(20, Ms), 2 * Ku / mu0, [None] * 3, tuple(3 * u))
    assert np.allclose(h2, 3 * h, rtol=1e-13, atol=0)
    assert np.all(po.uniaxial_anisotropy(m, None, np.zeros(20), 1.0, [None] * 3, tuple(u)) == 0)


def test_nonlocal_damping_fourier_eigenvalue():
    """PBC, uniform lambda_e: t = -lambda_e * Laplacian(H), a Fourier mode's eigenvalue."""
    N = (4, 6, 10)  # N1 != N2 exercises the axis-1 neighbour addressing
    cs = (2e-9, 3e-9, 1e-9)
    kk = (1, 2, 3)
    grid = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    ph = sum(2 * math.pi * kk[a] * grid[a] / N[a] for a in range(3))
    H = np.stack([np.cos(ph), np.sin(ph), np.cos(2 * ph)]) * 1e5
    lmul = (2e-19, 3e-19, 5e-19)
    t = po.llbar_nonlocal00nc(H, None, [None] * 3, lmul, N, cs, (1, 1, 1)).reshape((3,) + N)
    lam1 = sum(2 * (math.cos(2 * math.pi * kk[a] / N[a]) - 1) / cs[a] ** 2 for a in range(3))
    lam2 = sum(2 * (math.cos(4 * math.pi * kk[a] / N[a]) - 1) / cs[a] ** 2 for a in range(3))
    want = np.stack([-lmul[0] * lam1 * H[0], -lmul[1] * lam1 * H[1], -lmul[2] * lam2 * H[2]])
    assert np.abs(t - want).max() <= 1e-9 * np.abs(want).max()
    # uniform H: zero (open or periodic); masked cells: zero
    u = np.ones((3,) + N)
    assert np.abs(po.llbar_nonlocal00nc(u, None, [None] * 3, lmul, N, cs, (0, 0, 0))).max() < 1e-20
    msk = np.zeros(N)
    assert np.all(po.llbar_nonlocal00nc(H, msk, [None] * 3, lmul, N, cs, (1, 1, 1)) == 0)


def test_div_mul_pow_single_precision_power():
    rng = np.random.default_rng(5)
    a, b, c = rng.uniform(-1, 1, 100), rng.uniform(0.1, 2, 100), rng.uniform(0.01, 3, 100)
    b[3] = 0.0
    c[4] = 0.0
    got = po.div_mul_pow(a, b, c, 2.5)
    pw = np.power(c.astype(np.float32), np.float32(2.5)).astype(np.float64)
    want = np.where(b == 0, 0.0, a / np.where(b == 0, 1, b)) * np.where(c == 0, 0.0, pw)
    assert np.allclose(got, want, rtol=2e-7, atol=0)
    assert got[3] == 0.0 and got[4] == 0.0


def test_wavg_special_cases():
    rng = np.random.default_rng(6)
    x0, x1 = rng.uniform(100, 1000, 40), rng.uniform(100, 1000, 40)
    w0, w1 = rng.uniform(0.1, 2, 40), rng.uniform(0.1, 2, 40)
    # R = 0: the plain weighted mean
    got = po.wavg(x0, x1, w0, w1, None, 1.0, 1.0, 0.0)
    assert np.allclose(got, (w0 * x0 + w1 * x1) / (w0 + w1), rtol=1e-14, atol=0)
    # equal temperatures average to themselves for any correlation R
    assert np.allclose(po.wavg(x0, x0, w0, w1, None, 1.0, 1.0, 0.7), x0, rtol=1e-14, atol=0)
    # one weight zero: the other temperature
    assert np.allclose(po.wavg(x0, x1, w0, None, None, 1.0, 0.0, 0.7), x0, rtol=1e-14, atol=0)
    # R = 1: (sqrt(w0 x0) + sqrt(w1 x1))^2 / (sqrt(w0) + sqrt(w1))^2
    r1 = po.wavg(x0, x1, w0, w1, None, 1.0, 1.0, 1.0)
    want = (np.sqrt(w0 * x0) + np.sqrt(w1 * x1)) ** 2 / (np.sqrt(w0) + np.sqrt(w1)) ** 2
    assert np.allclose(r1, want, rtol=1e-13, atol=0)
    assert po.wavg(x0[:1], x1[:1], np.zeros(1), np.zeros(1), None, 1.0, 1.0, 0.5)[0] == 0.0
