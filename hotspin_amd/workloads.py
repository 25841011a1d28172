"""Synthetic LLB-r workloads on one MI355X: material parameters, initial
states and request streams.

Material of the LLB-r recipes (examples/p2f.py, examples/laser-demag.py of
the reference): msat0T0, lex = sqrt(2A/(mu0 Ms0^2)), J (spin), n = Ms0 /
(g muB J) with g = 2, Tc, T, lambda on all three damping axes, gamma_LL =
2.211e5.  Meshes are in internal (z, y, x) order: a user mesh X x Y x Z is
N = (Z, Y, X).

- `thin_film()`: the LLG limit of LLB-r on a 128x32x1 film of 3.90625 x
  3.90625 x 3 nm cells (Ms = 8e5 A/m, A = 1.3e-11 J/m, alpha = 0.02), m
  along +x tilted per cell by up to 10 degrees, an in-plane field of 20-30 mT
  at 150-190 degrees, fixed RK4 at dt = 1e-13 s, lazy normalisation.
- `single()`: one 512x512x16 solve at 300 K, 2 nm cubes, RK23 tol 1e-5.
- `laser()`: 256x256x8 film heated by a Gaussian pump, LLB-r coupled to the
  two-temperature model (T = Te), 5 ps of adaptive RK23.
- `ensemble()` (material MATERIAL_WIDE_T): 512 solves, seeds 0..511, meshes drawn per seed from
  {32x32x32, 64x64x4, 128x128x2, 96x48x8}, 3 nm cells, T ~ U(100, 1000) K,
  fixed RK4 at dt = 1e-15 s.
- `requests()` (material MATERIAL_WIDE_T): Poisson arrivals at 200 req/s; cell count log-uniform in
  [16^3, 256x256x4] mapped to the nearest power-of-two shape class; T ~
  U(0, 1.2 Tc) (clamped below at 0.1 Tc, where the mean-field longitudinal
  susceptibility vanishes like exp(-Tc/T) and the explicit LLB-r step has
  no finite stable dt); B_ext components ~ U(-100, 100) mT; 200..5000 RK23
  steps at tolerance 1e-5.
"""
import functools
import math

import numpy as np

MU0 = 4e-7 * math.pi
MUB = 9.27400968202020202e-24
GAMMA0 = 2.211e5

MATERIAL = dict(Ms0=1.7e6, Tc=1044.0, A=2.1e-11, lam=0.01, J=0.5)
# The ensemble and serving workloads reach down to 100 K.  With J = 1/2 the
# mean-field longitudinal susceptibility is so small there (exp(-2 Tc/T)-like)
# that explicit RK4 at dt = 1e-15 s diverges below ~170 K (FP64 oracle and
# GPU alike); J = 7/2 keeps the same Tc and stays stable down to 100 K.
MATERIAL_WIDE_T = dict(MATERIAL, J=3.5)


def equilibrium_m(T, Tc, J):
    """m_e(T): root of B_J(3J/(J+1) Tc/T m) = m by bisection (host set-up)."""
    if T <= 0:
        return 1.0
    if T >= Tc:
        return 0.0
    a, b = (2 * J + 1) / (2 * J), 1 / (2 * J)
    pre = 3 * J / (J + 1) * Tc / T

    def f(m):
        x = pre * m
        return a / math.tanh(a * x) - b / math.tanh(b * x) - m

    lo, hi = 1e-6, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def llbr_params(T=300.0, seed=1234, thermal=True, material=MATERIAL, demag=True, Bext=None):
    """Keyword arguments of api.Solver / api.Batch for the LLB-r composition
    exchange6 + demag + longitudinal field + Brown thermal field + llbar
    torque + local00nc damping."""
    Ms0, J = material["Ms0"], material["J"]
    kw = dict(msat0T0Mul=Ms0, exchange=1, lexMul=math.sqrt(2 * material["A"] / (MU0 * Ms0 * Ms0)),
              demag=1 if demag else 0, demag_accuracy=6, longfield=1, JMul=J, nMul=Ms0 / (2.0 * MUB * J),
              TcMul=material["Tc"], TMul=float(T), thermal=1 if thermal else 0, seed=int(seed), stream=0,
              cutoff_dt=0.0, gammaLL=GAMMA0, torque=1, local00nc=1, alphaMul=(material["lam"],) * 3)
    if Bext is not None:
        kw.update(zeeman=1, Bext=tuple(float(b) for b in Bext))
    return kw


def random_domains(N, block, me, rng):
    """Random unit direction per block^3 cells, scaled to |m| = me; 3 x ncell."""
    nb = [max(1, n // block) for n in N]
    dirs = rng.standard_normal((3,) + tuple(nb))
    dirs /= np.linalg.norm(dirs, axis=0)
    idx = [np.minimum(np.arange(n) // block, nb[a] - 1) for a, n in enumerate(N)]
    m = dirs[:, idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]] * me
    return np.ascontiguousarray(m.reshape(3, -1))


def random_state(ncell, me, rng):
    """Independent random direction per cell, |m| = me; 3 x ncell."""
    m = rng.standard_normal((3, ncell))
    m *= me / np.linalg.norm(m, axis=0)
    return m


def thin_film(seed=0):
    """128x32x1 LLG-limit film (seed -> numpy PCG64 stream: |B|, angle,
    then per-cell tilt and azimuth).  Returns (N, cell, kw, m0)."""
    rng = np.random.default_rng(seed)
    Ms, A, alpha = 8.0e5, 1.3e-11, 0.02
    bmag = rng.uniform(0.020, 0.030)
    phi = math.radians(rng.uniform(150.0, 190.0))
    N = (1, 32, 128)
    cell = (3e-9, 125e-9 / 32, 500e-9 / 128)
    ncell = N[0] * N[1] * N[2]
    theta = np.radians(rng.uniform(0.0, 10.0, ncell))
    psi = rng.uniform(0.0, 2 * math.pi, ncell)
    # user (x, y, z) -> internal (z, y, x)
    m = np.stack([np.sin(theta) * np.sin(psi), np.sin(theta) * np.cos(psi), np.cos(theta)])
    kw = dict(msat0T0Mul=Ms, exchange=1, lexMul=math.sqrt(2 * A / (MU0 * Ms * Ms)), demag=1, demag_accuracy=6,
              zeeman=1, Bext=(0.0, bmag * math.sin(phi), bmag * math.cos(phi)), gammaLL=GAMMA0 / (1 + alpha ** 2),
              torque=1, local00nc=1, alphaMul=(alpha,) * 3, normalize=1)
    return N, cell, kw, np.ascontiguousarray(m)


def single():
    """512x512x16 LLB-r solve at 300 K (2 nm cubes, padded 1024x1024x32)."""
    w = dict(name="llb-r T=300K 512x512x16 cells 2nm, demag padded 1024x1024x32, RK23 tol 1e-5",
             N=(16, 512, 512), cell=(2e-9, 2e-9, 2e-9), T=300.0, seed=1234, tol=1e-5, block=16)
    me = equilibrium_m(w["T"], MATERIAL["Tc"], MATERIAL["J"])
    m = random_domains(w["N"], w["block"], me, np.random.default_rng(w["seed"]))
    return w, llbr_params(T=w["T"], seed=w["seed"]), m, me


def two_temperature_peak(Q0, gam, Cl, G, T0, sig_t, t0, h=0.25e-15):
    """Peak Te of the spot centre: dTe/dt = (Q(t) - G (Te - Tp)) / (gam Te),
    dTp/dt = G (Te - Tp) / Cl, Q(t) = Q0 exp(-(t - t0)^2 / (2 sig_t^2)), from
    Te = Tp = T0 at t = 0, classical RK4 at step h up to t0 + 8 sig_t.  Lateral
    diffusion over that time is a few nm against the 200 nm spot, so the
    centre cell follows this ODE."""
    def f(t, te, tp):
        q = Q0 * math.exp(-(t - t0) ** 2 / (2.0 * sig_t * sig_t))
        x = G * (te - tp)
        return (q - x) / (gam * te), x / Cl

    te = tp = T0
    peak = T0
    n = int(math.ceil((t0 + 8 * sig_t) / h))
    for i in range(n):
        t = i * h
        a1, b1 = f(t, te, tp)
        a2, b2 = f(t + h / 2, te + h / 2 * a1, tp + h / 2 * b1)
        a3, b3 = f(t + h / 2, te + h / 2 * a2, tp + h / 2 * b2)
        a4, b4 = f(t + h, te + h * a3, tp + h * b3)
        te += h / 6 * (a1 + 2 * a2 + 2 * a3 + a4)
        tp += h / 6 * (b1 + 2 * b2 + 2 * b3 + b4)
        peak = max(peak, te)
    return peak


@functools.lru_cache(maxsize=None)
def pump_q0(gam, Cl, G, T0, Tmax, sig_t, t0):
    """Pump amplitude Q0 (W/m^3) whose pulse drives the spot centre's electron
    temperature to a peak of Tmax, electron-phonon losses during the pulse
    included (bisection on two_temperature_peak, which grows with Q0).  The
    lossless estimate gam (Tmax^2 - T0^2) / 2 = Q0 sig_t sqrt(2 pi) brackets
    it from below."""
    lo = gam * (Tmax * Tmax - T0 * T0) / (2.0 * sig_t * math.sqrt(2.0 * math.pi))
    hi = 2.0 * lo
    while two_temperature_peak(hi, gam, Cl, G, T0, sig_t, t0) < Tmax:
        hi *= 2.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if two_temperature_peak(mid, gam, Cl, G, T0, sig_t, t0) < Tmax:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def laser(Tc=MATERIAL["Tc"]):
    """Laser-heated LLB-r + two-temperature model on a 256x256x8 film of
    4x4x2 nm cells: Ce = gamma Te (gamma = 6e3 J/m^3K^2), Cl = 3.8e6 J/m^3K,
    G = 1.8e18 W/m^3K, kappa = 90 W/mK, Gaussian pump (sigma_r = 200 nm at the
    film centre, sigma_t = 100 fs, centred at t0 = 300 fs), T0 = 300 K.  Q0 is
    set (`pump_q0`) so that the electron temperature at the spot centre peaks
    at Tmax = 1.3 Tc.  Span 5 ps, RK23 with bounds 1e-5 (m) and 1e-3 K (Te,
    Tp).  Returns (w, kw, tm, m0)."""
    N = (8, 256, 256)
    cell = (2e-9, 4e-9, 4e-9)
    T0, gam = 300.0, 6e3
    Tmax = 1.3 * Tc
    sig_t = 100e-15
    Cl, G = 3.8e6, 1.8e18
    tm = dict(gammaE=gam, Cl=Cl, G=G, kappa=90.0, Q0=pump_q0(gam, Cl, G, T0, Tmax, sig_t, 3 * sig_t),
              sigma_r=200e-9, sigma_t=sig_t, t0=3 * sig_t, xc=N[2] * cell[2] / 2, yc=N[1] * cell[1] / 2,
              maxErrT=1e-3)
    me = equilibrium_m(T0, MATERIAL["Tc"], MATERIAL["J"])
    rng = np.random.default_rng(3)
    ncell = N[0] * N[1] * N[2]
    th = np.radians(rng.uniform(0.0, 5.0, ncell))
    ps = rng.uniform(0.0, 2 * math.pi, ncell)
    m0 = me * np.stack([np.sin(th) * np.sin(ps), np.sin(th) * np.cos(ps), np.cos(th)])  # user +x
    kw = llbr_params(T=T0, seed=77)
    w = dict(name="laser LLB-r + 2TM 256x256x8 cells 4x4x2nm, 5 ps, RK23", N=N, cell=cell, T0=T0, span=5e-12,
             tol=1e-5, Tmax=Tmax)
    return w, kw, tm, np.ascontiguousarray(m0)


ENSEMBLE_SHAPES = [(32, 32, 32), (4, 64, 64), (2, 128, 128), (8, 48, 96)]  # user 32^3, 64x64x4, 128x128x2, 96x48x8


def ensemble(nsolve=512, seed=0):
    """Per-solve (shape, T, seed) of the 512-solve stochastic ensemble."""
    rng = np.random.default_rng(seed)
    shape_idx = rng.integers(0, len(ENSEMBLE_SHAPES), nsolve)
    T = rng.uniform(100.0, 1000.0, nsolve)
    return [dict(N=ENSEMBLE_SHAPES[int(shape_idx[q])], T=float(T[q]), seed=q) for q in range(nsolve)]


ENSEMBLE_CELL = (3e-9, 3e-9, 3e-9)

# serving shape classes by cell count (user X x Y x Z -> internal (Z, Y, X))
SERVE_SHAPES = [(16, 16, 16), (8, 32, 32), (4, 64, 64), (8, 64, 64), (4, 128, 128), (8, 128, 128), (4, 256, 256)]
SERVE_CELL = (3e-9, 3e-9, 3e-9)


def requests(n=2000, rate=200.0, seed=0, Tc=MATERIAL["Tc"]):
    """The serving request stream: arrival time (s), shape class, T, B_ext,
    steps, seed."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    logc = rng.uniform(math.log2(4096), math.log2(262144), n)
    cls = np.clip(np.rint(logc - 12).astype(int), 0, len(SERVE_SHAPES) - 1)
    T = np.maximum(rng.uniform(0.0, 1.2 * Tc, n), 0.1 * Tc)
    B = rng.uniform(-0.1, 0.1, (n, 3))
    steps = rng.integers(200, 5001, n)
    return [dict(id=i, t=float(t[i]), cls=int(cls[i]), N=SERVE_SHAPES[int(cls[i])], T=float(T[i]),
                 Bext=tuple(B[i]), steps=int(steps[i]), seed=i) for i in range(n)]
