"""Pins of the FP64 oracle by mathematics: closed forms, identities, known-answer
vectors and brute force on tiny inputs (for the parts no reference vector covers)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

MU0 = 4e-7 * math.pi


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 25, 30, 49, 64, 96, 100, 128, 192, 1024, 11, 52])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_matches_numpy(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    buf = np.empty(2 * n)
    buf[0::2], buf[1::2] = x.real, x.imag
    assert po.lib().or_fft1d(buf.ctypes.data, n, sign) == 0
    got = buf[0::2] + 1j * buf[1::2]
    want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))


# ---------------------------------------------------------------- demag kernel
def _brute_conv(kern_real, P, N, m, msk, Ms):
    """Real-space circular convolution B = mu0 Ms sum K(r-r') M(r'), tiny meshes."""
    full = {}
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (1, 2): 3, (2, 1): 3, (0, 2): 4, (2, 0): 4, (0, 1): 5, (1, 0): 5}
    M = m * msk[None]
    B = np.zeros((3,) + N)
    for i in range(N[0]):
        for j in range(N[1]):
            for k in range(N[2]):
                for ii in range(N[0]):
                    for jj in range(N[1]):
                        for kk in range(N[2]):
                            di, dj, dk = (i - ii) % P[0], (j - jj) % P[1], (k - kk) % P[2]
                            for d in range(3):
                                for s in range(3):
                                    B[d, i, j, k] += kern_real[idx[(s, d)], di, dj, dk] * M[s, ii, jj, kk]
    return MU0 * Ms * B


@pytest.mark.parametrize("N,pbc", [((2, 3, 4), (0, 0, 0)), ((1, 4, 3), (0, 0, 0)), ((3, 2, 2), (0, 1, 0)),
                                   ((2, 3, 5), (0, 0, 1)), ((3, 4, 1), (0, 0, 0)), ((7, 2, 3), (0, 0, 0))])
def test_demag_fft_equals_real_space_convolution(N, pbc):
    """Includes an odd periodic x length (P2 = 5), a one-cell x extent (P2 = 1)
    and radix 7 (P0 = 14)."""
    cell = (2e-9, 3e-9, 2.5e-9)
    d = po.Demag(N, cell, pbc)
    kern = po.kernel_arne(d.P, cell, pbc)
    rng = np.random.default_rng(3)
    m = rng.standard_normal((3,) + N)
    msk = rng.random(N)
    B = d.B(m, msk, 7e5)
    Bb = _brute_conv(kern, d.P, N, m, msk, 7e5)
    assert np.max(np.abs(B - Bb)) <= 1e-12 * np.max(np.abs(Bb))


def test_kernel_cube_self_term_and_far_field():
    """Cube cell: N_xx=N_yy=N_zz=1/3 (kernel = -N); far field -> point dipole."""
    c = 3e-9
    P = (48, 48, 48)
    K = po.kernel_arne(P, (c, c, c), (0, 0, 0))
    for comp in range(3):
        assert abs(K[comp, 0, 0, 0] + 1.0 / 3.0) < 2e-3
    assert abs(K[0, 0, 0, 0] - K[1, 0, 0, 0]) < 1e-12
    assert abs(K[3:, 0, 0, 0]).max() < 1e-12
    # dipole at r = (0, 0, 20) cells along internal axis 2: K_22 = 2V/(4 pi r^3), K_00 = -V/(4 pi r^3)
    r = 20 * c
    V = c ** 3
    assert abs(K[2, 0, 0, 20] / (2 * V / (4 * math.pi * r ** 3)) - 1) < 2e-3
    assert abs(K[0, 0, 0, 20] / (-V / (4 * math.pi * r ** 3)) - 1) < 2e-3
    # generic off-axis point: XY entry 3 x y V / (4 pi r^5)
    x, y, z = 7 * c, 9 * c, 11 * c
    r = math.sqrt(x * x + y * y + z * z)
    want = 3 * x * y * V / (4 * math.pi * r ** 5)
    assert abs(K[5, 7, 9, 11] / want - 1) < 2e-3


def test_kernel_symmetries():
    P = (8, 12, 10)
    K = po.kernel_arne(P, (1e-9, 2e-9, 1.5e-9), (0, 0, 0))
    XX, YY, ZZ, YZ, XZ, XY = K
    for i in range(1, 4):
        assert np.allclose(XX[P[0] - i], XX[i], rtol=0, atol=0)
        assert np.allclose(XY[P[0] - i], -XY[i], rtol=0, atol=0)
    # odd components vanish on the symmetry planes
    tol = 1e-12 * np.abs(K).max()
    assert np.abs(XY[0]).max() < tol and np.abs(XZ[:, :, 0]).max() < tol and np.abs(YZ[:, 0]).max() < tol


def test_thin_film_limit():
    """Periodic thin film magnetised out of plane: B_normal -> -mu0 M."""
    N = (1, 8, 8)
    d = po.Demag(N, (2e-9, 2e-9, 2e-9), (0, 32, 32))
    m = np.zeros((3,) + N)
    m[0] = 1.0
    B = d.B(m, None, 8e5)
    assert abs(B[0].mean() / (-MU0 * 8e5) - 1) < 1e-2
    assert np.abs(B[1:]).max() < 1e-3 * MU0 * 8e5


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, out):
    """Random123 known-answer vectors for philox4x32-10."""
    assert list(po.philox4x32_10(ctr, key)) == out


def test_thermal_normals_moments_and_determinism():
    n = 200000
    a = po.thermal_normals(1234, 0, 7, 0, n)
    b = po.thermal_normals(1234, 0, 7, 0, n)
    c = po.thermal_normals(1234, 0, 8, 0, n)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    x = a.ravel()
    se = 1 / math.sqrt(x.size)
    assert abs(x.mean()) < 5 * se
    assert abs(x.var() - 1) < 5 * math.sqrt(2) * se
    assert abs(np.mean(x ** 4) - 3) < 0.05
    assert abs(np.corrcoef(a[0], a[1])[0, 1]) < 5 * se * math.sqrt(3)


# ---------------------------------------------------------------- MFA
def _lf(m, S, T, Tc=1044.0, n=1e28, Ms=1.7e6):
    N = m.shape[1]
    h = np.empty_like(m)
    po.lib().or_long_field(h[0].ctypes.data, h[1].ctypes.data, h[2].ctypes.data,
                           m[0].ctypes.data, m[1].ctypes.data, m[2].ctypes.data,
                           None, None, None, None, None, Ms, S, n, Tc, T, N)
    return h


def _me(T, S, Tc=1044.0):
    N = len(T)
    out = np.empty(N)
    Tarr = np.ascontiguousarray(T, dtype=np.float64)
    po.lib().or_brillouin(out.ctypes.data, None, Tarr.ctypes.data, None, None, 1.0, 1.0, Tc, S, N)
    return out


@pytest.mark.parametrize("S", [0.5, 1.0, 3.5, 2e5])
def test_brillouin_root_is_a_fixed_point(S):
    T = np.array([1.0, 100.0, 300.0, 700.0, 1000.0, 1043.0, 1044.0, 1100.0])
    me = _me(T, S)
    assert me[-1] == 0.0
    for t, x in zip(T[:-1], me[:-1]):
        pre = 3 * (1.0 if S > 1e5 else S / (S + 1)) * 1044.0 / t
        if S > 1e5:
            val = 1 / math.tanh(pre * x) - 1 / (pre * x) if x > 0 else 0
        else:
            a, b = (2 * S + 1) / (2 * S), 1 / (2 * S)
            val = a / math.tanh(a * pre * x) - b / math.tanh(b * pre * x) if x > 0 else 0
        assert abs(val - x) < 1e-7
    if S > 1e5:
        assert abs((1 - me[0]) / (1.0 / (3 * 1044.0)) - 1) < 1e-3  # L(x) ~ 1 - 1/x
    else:
        assert abs(me[0] - 1) < 1e-6
    assert np.all(np.diff(me[:-1]) <= 0)


def test_long_field_vanishes_at_equilibrium_and_restores():
    S = 0.5
    T = 600.0
    me = _me(np.array([T]), S)[0]
    rng = np.random.default_rng(0)
    u = rng.standard_normal((3, 16))
    u /= np.linalg.norm(u, axis=0)
    for scale, sign in [(1.0, 0), (0.9, 1), (1.1, -1)]:
        m = np.ascontiguousarray(u * me * scale)
        h = _lf(m, S, T)
        proj = (h * u).sum(axis=0)
        if sign == 0:
            assert np.abs(h).max() < 1e-6 * np.abs(_lf(np.ascontiguousarray(u * me * 0.9), S, T)).max()
        else:
            assert np.all(np.sign(proj) == sign)  # pushes |m| back towards m_e
        # parallel to m
        assert np.abs(np.cross(h.T, u.T)).max() <= 1e-12 * max(1.0, np.abs(h).max())


def test_brillouin_series_branch_continuity():
    """B_S and dB_S/dx switch to series below x = 0.2/((2S+1)/(2S)); both sides agree."""
    S = 0.5
    gpre = (2 * S + 1) / (2 * S)
    lim = 0.2 / gpre
    Tc, T = 1044.0, 500.0
    b = 3 * S / (S + 1) * Tc / T
    for x in (lim * (1 - 1e-9), lim * (1 + 1e-9)):
        mabs = x / b
        m = np.ascontiguousarray(np.array([[mabs], [0.0], [0.0]]))
        h = _lf(m, S, T)[0, 0]
        # B_1/2(x) = tanh(x); dB/dx = 1 - tanh^2
        W, dW = math.tanh(x), 1 - math.tanh(x) ** 2
        want = 1e28 * 1.38064881313131313e-23 * T * (W - mabs) / (MU0 * 1.7e6 * dW)
        assert abs(h / want - 1) < 1e-8  # series truncated after x^7 (common_func.h:89-92)


# ---------------------------------------------------------------- solvers
def _macrospin(alpha=0.0, normalize=0):
    B = (0.0, 0.0, 0.1)  # tesla along internal axis 2
    return po.LLB((1, 1, 1), (2e-9, 2e-9, 2e-9), zeeman=1, Bext=B, gammaLL=2.211e5, torque=1,
                  local00nc=1 if alpha else 0, alphaMul=(alpha, alpha, alpha), msat0T0Mul=8e5,
                  normalize=normalize)


def _exact_precession(m0, t):
    """dm/dt = gamma m x H (internal frame), H along axis 2."""
    w = 2.211e5 * 0.1 / MU0
    c, s = math.cos(w * t), math.sin(w * t)
    return np.array([c * m0[0] + s * m0[1], -s * m0[0] + c * m0[1], m0[2]])


@pytest.mark.parametrize("method,order", [("euler", 1), ("heun", 2), ("rk4", 4)])
def test_fixed_step_orders(method, order):
    s = _macrospin()
    m0 = np.array([0.6, 0.0, 0.8])
    T = 2e-11
    errs = []
    for nst in (40, 80):
        m = np.ascontiguousarray(m0.reshape(3, 1).copy())
        dt = T / nst
        for i in range(nst):
            s.step(method, m, dt, i)
        errs.append(np.abs(m[:, 0] - _exact_precession(m0, T)).max())
    rate = math.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.25, (errs, rate)


def test_rk23_local_order_and_controller():
    s = _macrospin()
    m0 = np.array([0.6, 0.0, 0.8])
    errs = []
    for h in (2e-13, 1e-13):
        m = np.ascontiguousarray(m0.reshape(3, 1).copy())
        dt, st = s.step("rk23", m, h, 0, maxErr=1.0)
        errs.append(np.abs(m[:, 0] - _exact_precession(m0, h)).max())
        assert dt == pytest.approx(h * 1.5)  # tiny error: factor clamped to 1.5
    rate = math.log2(errs[0] / errs[1])
    assert abs(rate - 4) < 0.3  # local error of a 3rd-order method
    # a tolerance far below the step error forces rejections
    m = np.ascontiguousarray(m0.reshape(3, 1).copy())
    dt, st = s.step("rk23", m, 1e-11, 0, maxErr=1e-9)
    assert st[1] > 1 and st[0] <= 1e-9 and dt < 1e-11


def test_rk12_rejects_and_matches_heun_when_accepted():
    s = _macrospin(alpha=0.05, normalize=1)
    m0 = np.array([0.6, 0.0, 0.8])
    m1 = np.ascontiguousarray(m0.reshape(3, 1).copy())
    m2 = m1.copy()
    dt, st = s.step("rk12", m1, 1e-13, 0, maxErr=1.0)
    assert st[1] == 1
    s.step("heun", m2, 1e-13, 0)
    assert np.array_equal(m1, m2)  # accepted RK12 step == Heun step (solve_rk12.go vs solve_heun.go)
    m3 = np.ascontiguousarray(m0.reshape(3, 1).copy())
    dt, st = s.step("rk12", m3, 1e-10, 0, maxErr=1e-8, mindt=1e-18)
    assert st[1] > 1 and dt < 1e-10


def test_llg_limit_damping_relaxes_towards_field():
    s = _macrospin(alpha=0.5, normalize=1)
    m = np.ascontiguousarray(np.array([[0.6], [0.0], [0.8]]))
    for i in range(4000):
        s.step("rk4", m, 2e-13, i)
    n = m[:, 0] / np.linalg.norm(m[:, 0])
    assert n[2] > 0.999


def test_torque_is_orthogonal_and_zero_for_parallel():
    rng = np.random.default_rng(5)
    n = 64
    m = rng.standard_normal((3, n))
    H = rng.standard_normal((3, n))
    t = np.empty_like(m)
    L = po.lib()
    L.or_llbar_torque(t[0].ctypes.data, t[1].ctypes.data, t[2].ctypes.data, m[0].ctypes.data,
                      m[1].ctypes.data, m[2].ctypes.data, H[0].ctypes.data, H[1].ctypes.data,
                      H[2].ctypes.data, None, n)
    assert np.abs((t * m).sum(0)).max() < 1e-12
    assert np.allclose(t, np.cross(m.T, H.T).T, atol=1e-14)  # M x H in the internal frame
    Hp = np.ascontiguousarray(m * 3.0)
    L.or_llbar_torque(t[0].ctypes.data, t[1].ctypes.data, t[2].ctypes.data, m[0].ctypes.data,
                      m[1].ctypes.data, m[2].ctypes.data, Hp[0].ctypes.data, Hp[1].ctypes.data,
                      Hp[2].ctypes.data, None, n)
    assert np.abs(t).max() < 1e-12


# ---------------------------------------------------------------- exchange / heat stencils
def test_exchange_fourier_mode_eigenvalue():
    """PBC, uniform lex: a Fourier mode is an eigenvector of the discrete Laplacian."""
    N = (4, 8, 16)
    cs = (2e-9, 3e-9, 1e-9)
    kk = (1, 2, 3)
    grid = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    ph = sum(2 * math.pi * kk[a] * grid[a] / N[a] for a in range(3))
    m = np.stack([np.cos(ph), np.sin(ph), np.zeros(N)])
    lex = 5e-9
    pre = [lex * lex / c ** 2 for c in cs]
    H = po.exchange6(m, None, N, (1, 1, 1), 8e5, pre).reshape((3,) + N)
    lam = sum(2 * (math.cos(2 * math.pi * kk[a] / N[a]) - 1) / cs[a] ** 2 for a in range(3))
    want = 8e5 * lex * lex * lam * m
    assert np.abs(H - want).max() <= 1e-9 * np.abs(want).max()
    # uniform m: zero field; open boundaries: zero too (Neumann)
    mu = np.ones((3,) + N)
    assert np.abs(po.exchange6(mu, None, N, (0, 0, 0), 8e5, pre)).max() == 0


def test_qspat_conserves_and_diffuses():
    N = (2, 6, 9)
    rng = np.random.default_rng(2)
    T = 300 + 50 * rng.random(N)
    k = 1 + rng.random(N)
    Q = np.empty(N)
    for pbc in [(1, 1, 1), (0, 0, 0)]:
        po.lib().or_qspat(Q.ctypes.data, T.ctypes.data, k.ctypes.data, 90.0, N[0], N[1], N[2],
                          2e-9, 4e-9, 4e-9, *pbc)
        # harmonic-mean conductances are symmetric: total heat is conserved
        assert abs(Q.sum()) <= 1e-10 * np.abs(Q).sum()
    # hottest cell cools, coldest warms
    assert Q.flat[np.argmax(T)] < 0 and Q.flat[np.argmin(T)] > 0
