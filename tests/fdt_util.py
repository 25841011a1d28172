"""Fluctuation-dissipation checks of the Brown thermal field (shared by the
oracle and GPU tests).

The reference adds H_therm to H_eff (examples/p2f.py:107-114, mu = alpha)
with per-component standard deviation
    sigma = sqrt(2 kB T lambda / (mu0 V gamma_LL dt' Ms0)),  dt' = max(cutoff_dt, dt)
(modules/mod_brown.go:124-140, src/libhotspin/temperature.cu:14-71), and the
LLB-r right-hand side is dm/dt = gamma_LL (H x m + lambda H)
(llbar-torque.cu:11-42, llbar-local00nc.cu:11-60).  For non-interacting
cells (no exchange, no demag) in a uniform field B along x this SDE has a
closed-form stationary state (Fokker-Planck, Stratonovich; the solver keeps
the noise fixed over a step, which converges to that interpretation):

* Transverse (angle theta between m and B): the precession noise gamma H_th x m
  rotates m with angular diffusion gamma lambda kB T / (mu0 V Ms0) and the
  damping noise adds lambda^2/|m|^2 of it; the damping drift is
  gamma lambda H_perp / |m|.  So P(theta) ~ exp(xi cos theta) with
      xi = Ms0 V B |m| / (kB T (|m|^2 + lambda^2)),
  and <cos theta> = L(xi) = coth xi - 1/xi (Langevin).  In the LLG limit
  (normaliser, |m| = 1) this is Brown's result at temperature T (1 + lambda^2).
  With a free |m| (LLB-r) the field also pushes r by gamma lambda B cos(theta),
  with the radial noise lambda^2 kB T instead of (r^2 + lambda^2) kB T: the two
  zero-flux conditions disagree, the stationary state carries a circulating
  flux, and L(xi(<r>)) is exact only where the field barely moves r.  Near
  Tc (large longitudinal susceptibility) a 1.3 T field biases <cos theta> by
  about +1 % (oracle and GPU alike, 1000 K); 0.3 T keeps the bias below the
  statistical error of 32768 cells.
* Radial (LLB-r, B = 0): only the additive damping noise gamma lambda H_th
  moves |m| = r, so P(r) ~ r^2 exp(-E(r) / (lambda^2 kB T)) with
  E(r) = -mu0 Ms0 V int H_lf dr: <r> is the MFA equilibrium m_e(T) up to
  O(lambda^2) and Var(r) is proportional to the noise amplitude squared.

A wrong prefactor (a factor 2, a wrong volume, dt instead of
max(cutoff_dt, dt)) moves xi or Var(r) by tens of standard errors.  Standard
errors come from per-cell time averages: the cells are independent, so the
spread of their averages over cells is an honest error bar.
"""
import math

import numpy as np

KB = 1.38064881313131313e-23
MU0 = 4e-7 * math.pi
MUB = 9.27400968202020202e-24
GAMMA = 2.211e5
MS = 8e5
TC = 1044.0


def langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


def m_e_half(T, Tc=TC):
    """MFA equilibrium of spin 1/2 (B_1/2 = tanh): root of m = tanh(Tc/T m), bisection."""
    lo, hi = 1e-9, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.tanh(Tc / T * mid) > mid:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def params(T, B, lam=0.1, normalize=1, longfield=0, cutoff_dt=0.0, seed=7):
    """Macrospin cells: Zeeman + (long field) + Brown, torque and local00nc on."""
    J = 0.5
    return dict(msat0T0Mul=MS, exchange=0, demag=0, zeeman=1, Bext=(B, 0.0, 0.0), longfield=longfield, JMul=J,
                nMul=MS / (2 * MUB * J), TcMul=TC, TMul=T, thermal=1, seed=seed, cutoff_dt=cutoff_dt,
                gammaLL=GAMMA, torque=1, local00nc=1, alphaMul=(lam,) * 3, normalize=normalize)


class Sampler:
    """Per-cell time averages of cos(theta) (angle to +x), |m| and |m|^2."""

    def __init__(self, ncell):
        self.c = np.zeros(ncell)
        self.r = np.zeros(ncell)
        self.r2 = np.zeros(ncell)
        self.n = 0

    def add(self, m):
        r = np.sqrt((m * m).sum(axis=0))
        self.c += m[0] / r
        self.r += r
        self.r2 += r * r
        self.n += 1

    def cos(self):
        """(mean over cells of the per-cell average, its standard error)"""
        a = self.c / self.n
        return a.mean(), a.std(ddof=1) / math.sqrt(a.size)

    def radius(self):
        a = self.r / self.n
        return a.mean(), a.std(ddof=1) / math.sqrt(a.size)

    def radial_var(self):
        """mean over cells of the per-cell time variance of |m|"""
        a = self.r / self.n
        return float((self.r2 / self.n - a * a).mean())


def run(advance, get_m, ncell, nequil, nsample, every=10):
    """advance(i0, n) takes steps i0 .. i0+n-1; nequil steps, then nsample
    steps sampled through get_m() every `every` steps."""
    s = Sampler(ncell)
    advance(0, nequil)
    for i in range(nequil, nequil + nsample, every):
        advance(i, every)
        s.add(get_m())
    return s


def xi_transverse(T, B, cell, lam, r=1.0, cutoff_dt=0.0, dt=1.0):
    """xi of the stationary angular density (r = |m|; the noise variance
    scales as dt / max(cutoff_dt, dt))."""
    V = cell[0] * cell[1] * cell[2]
    Teff = T * dt / max(cutoff_dt, dt)
    return MS * V * B * r / (KB * Teff * (r * r + lam * lam))


def radial_prediction(T, lam, cell, h_lf):
    """<r>, Var(r) of P(r) ~ r^2 exp(-E(r)/(lam^2 kB T)), E(r) = -mu0 Ms V int H_lf dr,
    by quadrature on a fine grid around m_e; h_lf(r) -> H_lf (A/m) along m."""
    me = m_e_half(T)
    r = np.linspace(max(me - 0.15, 1e-3), me + 0.15, 12001)
    H = h_lf(r)
    V = cell[0] * cell[1] * cell[2]
    E = np.concatenate([[0.0], np.cumsum(-MU0 * MS * V * 0.5 * (H[1:] + H[:-1]) * np.diff(r))])
    w = r * r * np.exp(-(E - E.min()) / (lam * lam * KB * T))
    w /= np.trapezoid(w, r)
    mean = np.trapezoid(r * w, r)
    var = np.trapezoid((r - mean) ** 2 * w, r)
    # the density must have decayed inside the grid
    assert w[0] < 1e-9 * w.max() and w[-1] < 1e-9 * w.max()
    return float(mean), float(var), me
