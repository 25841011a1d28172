"""Legacy FP64 libhotspin ABI on the GPU vs the FP64 oracle (same inputs).

Tolerance: relative 1e-12 of the output scale (FP64 with FMA contraction and
different transcendental implementations); reductions to 1e-12 (sums) and
bit-exact (maxima)."""
import math

import numpy as np
import pytest

from hotspin_amd._lib import check, lib
from oracle import pyoracle as po
from tests.gpu_util import P, dev, empty, host, relerr, require_gpu, torch  # noqa: F401

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("legacy")


@pytest.mark.parametrize("N,pbc,use_lex", [((3, 5, 70), (0, 0, 0), True), ((4, 9, 130), (1, 1, 1), True),
                                           ((1, 7, 64), (0, 1, 0), False), ((2, 3, 1), (1, 0, 1), True),
                                           ((5, 64, 64), (0, 0, 1), False)])
def test_exchange6(N, pbc, use_lex):
    rng = np.random.default_rng(sum(N))
    n = N[0] * N[1] * N[2]
    m = rng.standard_normal((3, n))
    lex = (0.5 + rng.random(n)) if use_lex else None
    pre = (11.0, 23.0, 7.0)
    want = po.exchange6(m, lex, N, pbc, 8e5, pre)
    dm, dl = dev(m), (dev(lex) if use_lex else None)
    h = empty((3, n))
    lib().exchange6Async(P(h[0]), P(h[1]), P(h[2]), P(dm[0]), P(dm[1]), P(dm[2]), None, P(dl),
                         N[0], N[1], N[2], *pbc, 8e5, *pre, None)
    assert relerr(host(h), want) <= TOL


def _rand(shape, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, shape)


def test_llbar_terms():
    n = 1000
    m, H = _rand((3, n), 1), _rand((3, n), 2) * 1e5
    msk = _rand(n, 3, 0, 1)
    msk[::7] = 0.0
    lam = [_rand(n, 4 + c, 0, 1) for c in range(3)]
    dm, dH, dmsk = dev(m), dev(H), dev(msk)
    dl = [dev(x) for x in lam]
    L = lib()
    t = empty((3, n))
    # torque
    want = np.empty((3, n))
    po.lib().or_llbar_torque(*[want[c].ctypes.data for c in range(3)], *[m[c].ctypes.data for c in range(3)],
                             *[H[c].ctypes.data for c in range(3)], msk.ctypes.data, n)
    L.llbar_torque_async(P(t[0]), P(t[1]), P(t[2]), P(dm[0]), P(dm[1]), P(dm[2]), P(dH[0]), P(dH[1]), P(dH[2]),
                         P(dmsk), None, n)
    assert relerr(host(t), want) <= TOL
    # local00nc with masks and with NULL masks
    for masks in (lam, [None, None, None]):
        dmk = [dev(x) if x is not None else None for x in masks]
        po.lib().or_llbar_local00nc(*[want[c].ctypes.data for c in range(3)], *[H[c].ctypes.data for c in range(3)],
                                    msk.ctypes.data, *[x.ctypes.data if x is not None else None for x in masks],
                                    0.1, 0.2, 0.3, n)
        L.llbar_local00nc_async(P(t[0]), P(t[1]), P(t[2]), P(dH[0]), P(dH[1]), P(dH[2]), P(dmsk),
                                *[P(x) for x in dmk], 0.1, 0.2, 0.3, None, n)
        assert relerr(host(t), want) <= TOL
    for oname, gname in (("or_llbar_local02nc", "llbar_local02nc_async"), ("or_llbar_local02c", "llbar_local02c_async")):
        getattr(po.lib(), oname)(*[want[c].ctypes.data for c in range(3)], *[m[c].ctypes.data for c in range(3)],
                                 *[H[c].ctypes.data for c in range(3)], msk.ctypes.data,
                                 *[x.ctypes.data for x in lam], 0.1, 0.2, 0.3, n)
        getattr(L, gname)(P(t[0]), P(t[1]), P(t[2]), P(dm[0]), P(dm[1]), P(dm[2]), P(dH[0]), P(dH[1]), P(dH[2]),
                          P(dmsk), *[P(x) for x in dl], 0.1, 0.2, 0.3, None, n)
        assert relerr(host(t), want) <= TOL


@pytest.mark.parametrize("S", [0.5, 3.5, 2e5])
def test_long_field_and_brillouin(S):
    n = 4096
    rng = np.random.default_rng(7)
    m = rng.standard_normal((3, n))
    m *= rng.uniform(0.01, 1.0, n) / np.linalg.norm(m, axis=0)  # physical |m| <= 1
    m[:, :3] = 0.0  # |m| <= 1e-32 branch
    m[:, 5] *= 1e-3  # series branch
    # T >= 100 K: towards T -> 0, dB/dx underflows to 0 and the field is inf on either side
    Tm = rng.uniform(100.0, 1000.0, n)
    Tm[:4] = 0.0
    msk = rng.uniform(0, 1, n)
    msk[10] = 0.0
    want = np.empty((3, n))
    Ts = np.maximum(Tm, 100.0)
    po.lib().or_long_field(*[want[c].ctypes.data for c in range(3)], *[m[c].ctypes.data for c in range(3)],
                           msk.ctypes.data, None, None, None, Ts.ctypes.data, 1.7e6, S, 1e28, 1044.0, 1.0, n)
    dm, dmsk, dT = dev(m), dev(msk), dev(Ts)
    h = empty((3, n))
    lib().long_field_async(P(h[0]), P(h[1]), P(h[2]), P(dm[0]), P(dm[1]), P(dm[2]), P(dmsk), None, None, None,
                           P(dT), 1.7e6, S, 1e28, 1044.0, 1.0, n, None)
    # dB_S/dx = (g^2-l^2) + l^2 coth^2(lx) - g^2 coth^2(gx) cancels (common_func.h:95-98):
    # ulp-level tanh differences are amplified by |coth^2|/|dW| at low T
    assert relerr(host(h), want) <= 1e-6
    # brillouin: Ridders iterations amplify rounding differences only to ~eps
    wm = np.empty(n)
    po.lib().or_brillouin(wm.ctypes.data, msk.ctypes.data, Tm.ctypes.data, None, None, 1.0, 1.7e6, 1044.0, S, n)
    dTm = dev(Tm)
    g = empty(n)
    lib().brillouinAsync(P(g), P(dmsk), P(dTm), None, None, 1.0, 1.7e6, 1044.0, S, n, None)
    assert relerr(host(g), wm) <= 1e-8


def test_temperature_scale_and_normalize():
    n = 2000
    h0 = _rand((3, n), 11)
    mu = [_rand(n, 12 + c, 0, 1) for c in range(3)]
    T = _rand(n, 20, 0, 900)
    msk = _rand(n, 21, 0, 1)
    msk[3] = 0
    want = h0.copy()
    po.lib().or_temperature_scale_noise(*[want[c].ctypes.data for c in range(3)], *[x.ctypes.data for x in mu],
                                        T.ctypes.data, msk.ctypes.data, 0.1, 0.2, 0.3, 7e12, n)
    dh = dev(h0)
    dmu = [dev(x) for x in mu]
    dT, dmsk = dev(T), dev(msk)  # keep the device buffers alive across the launch
    lib().temperature_scaleAnizNoise(P(dh[0]), P(dh[1]), P(dh[2]), *[P(x) for x in dmu], P(dT), P(dmsk),
                                     0.1, 0.2, 0.3, 7e12, None, n)
    assert relerr(host(dh), want) <= TOL
    m = _rand((3, n), 30)
    m[:, 7] = 0
    want = m.copy()
    po.lib().or_normalize(*[want[c].ctypes.data for c in range(3)], n)
    dm = dev(m)
    lib().normalizeAsync(P(dm[0]), P(dm[1]), P(dm[2]), None, n)
    assert relerr(host(dm), want) <= 1e-15


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1), (1, 0, 1)])
def test_qspat_qinter(pbc):
    N = (3, 17, 70)
    n = N[0] * N[1] * N[2]
    T = _rand(n, 40, 250, 900)
    k = _rand(n, 41, 0.5, 2)
    want = np.empty(n)
    po.lib().or_qspat(want.ctypes.data, T.ctypes.data, k.ctypes.data, 90.0, *N, 2e-9, 4e-9, 4e-9, *pbc)
    q = empty(n)
    dT, dk = dev(T), dev(k)
    lib().Qspat_async(P(q), P(dT), P(dk), 90.0, *N, 2e-9, 4e-9, 4e-9, *pbc, None)
    assert relerr(host(q), want) <= TOL
    T2 = _rand(n, 42, 250, 900)
    po.lib().or_qinter(want.ctypes.data, T.ctypes.data, T2.ctypes.data, k.ctypes.data, 1.8e18, n)
    dT2 = dev(T2)
    lib().QinterAsync(P(q), P(dT), P(dT2), P(dk), 1.8e18, n, None)
    assert relerr(host(q), want) <= TOL


def test_linear_algebra_family():
    n = 3001
    a, b, c = _rand(n, 50), _rand(n, 51), _rand(n, 52)
    b[5] = 0.0
    da, db, dc = dev(a), dev(b), dev(c)
    L = lib()
    d = empty(n)
    cases = [
        (lambda: L.addAsync(P(d), P(da), P(db), None, n), a + b),
        (lambda: L.linearCombination2Async(P(d), P(da), 0.3, P(db), -1.7, None, n), 0.3 * a - 1.7 * b),
        (lambda: L.linearCombination3Async(P(d), P(da), 0.3, P(db), -1.7, P(dc), 2.5, None, n),
         0.3 * a - 1.7 * b + 2.5 * c),
        (lambda: L.addMaddAsync(P(d), P(da), P(db), P(dc), 0.7, None, n), a + 0.7 * (b + c)),
        (lambda: L.maddAsync(P(d), P(da), P(db), 1.3, None, n), a + 1.3 * b),
        (lambda: L.maddAsync(P(d), P(da), None, 1.3, None, n), a + 1.3),
        (lambda: L.mulAsync(P(d), P(da), P(db), None, n), a * b),
        (lambda: L.divAsync(P(d), P(da), P(db), None, n), np.where(b == 0, 0, a / np.where(b == 0, 1, b))),
        (lambda: L.zeroArrayAsync(P(d), n, None), np.zeros(n)),
    ]
    for fn, want in cases:
        fn()
        assert relerr(host(d), want) <= TOL
    x = dev(a.copy())
    L.madd1Async(P(x), P(db), 0.5, None, n)
    assert relerr(host(x), a + 0.5 * b) <= TOL
    x = dev(a.copy())
    L.madd2Async(P(x), P(db), 0.5, P(dc), -0.25, None, n)
    assert relerr(host(x), a + (0.5 * b - 0.25 * c)) <= TOL
    x = dev(a.copy())
    L.madd2Async(P(x), None, 0.5, P(dc), -0.25, None, n)
    assert relerr(host(x), a + (0.5 - 0.25 * c)) <= TOL
    # vecMadd with and without b
    A = np.stack([a, b, c])
    B = np.stack([c, a, b])
    dA, dB = dev(A), dev(B)
    D = empty((3, n))
    L.vecMaddAsync(P(D[0]), P(D[1]), P(D[2]), P(dA[0]), P(dA[1]), P(dA[2]), P(dB[0]), P(dB[1]), P(dB[2]),
                   1.0, 2.0, 3.0, None, n)
    assert relerr(host(D), A + np.array([1.0, 2.0, 3.0])[:, None] * B) <= TOL
    L.vecMaddAsync(P(D[0]), P(D[1]), P(D[2]), P(dA[0]), P(dA[1]), P(dA[2]), None, None, None, 1.0, 2.0, 3.0, None, n)
    assert relerr(host(D), A + np.array([1.0, 2.0, 3.0])[:, None]) <= TOL
    # dot family
    L.dotMaskAsync(P(d), P(dA[0]), None, P(dA[2]), P(dB[0]), P(dB[1]), None, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, None, n)
    want = (1 * A[0]) * (4 * B[0]) + 2.0 * (5 * B[1]) + (3 * A[2]) * 6.0
    assert relerr(host(d), want) <= TOL
    L.dotAsync(P(d), P(dA[0]), P(dA[1]), P(dA[2]), P(dB[0]), P(dB[1]), P(dB[2]), None, n)
    assert relerr(host(d), (A * B).sum(0)) <= TOL
    Cc = np.stack([b, c, a])
    dC = dev(Cc)
    L.dotSignAsync(P(d), P(dA[0]), P(dA[1]), P(dA[2]), P(dB[0]), P(dB[1]), P(dB[2]), P(dC[0]), P(dC[1]), P(dC[2]),
                   None, n)
    assert relerr(host(d), np.copysign((A * B).sum(0), -(B * Cc).sum(0))) <= TOL
    # complex madd
    src = _rand(2 * n, 60)
    kern = _rand(n, 61)
    dst = _rand(2 * n, 62)
    ddst, dkern, dsrc = dev(dst), dev(kern), dev(src)
    L.cmaddAsync(P(ddst), 0.3, -0.8, P(dkern), P(dsrc), None, n)
    z = dst[0::2] + 1j * dst[1::2] + kern * (0.3 - 0.8j) * (src[0::2] + 1j * src[1::2])
    got = host(ddst)
    assert relerr(got[0::2] + 1j * got[1::2], z) <= TOL


def test_tens_symm_vec_mul_and_copypad():
    Nx, Ny, Nz = 8, 12, 9
    n = Nx * Ny * Nz
    src = [_rand(2 * n, 70 + c) for c in range(3)]
    K = [_rand((Nx // 2 + 1) * (Ny // 2 + 1) * Nz, 80 + c) for c in range(6)]
    want = [np.empty(2 * n) for _ in range(3)]
    po.lib().or_tens_symm_vec_mul(*[w.ctypes.data for w in want], *[s.ctypes.data for s in src],
                                  *[k.ctypes.data for k in K], 1.7, Nx, Ny, Nz)
    ds = [dev(s) for s in src]
    dk = [dev(k) for k in K]
    lib().tensSYMMVecMul(P(ds[0]), P(ds[1]), P(ds[2]), P(ds[0]), P(ds[1]), P(ds[2]), *[P(k) for k in dk], 1.7,
                         Nx, Ny, Nz, None)
    for c in range(3):
        assert relerr(host(ds[c]), want[c]) <= TOL
    # pad / unpad round trip, 2 components
    S = (3, 4, 5)
    D = (6, 4, 10)
    x = _rand((2,) + S, 90)
    dx = dev(x)
    dpad = dev(np.full((2,) + D, 7.0))
    lib().copyPad3DAsync(P(dpad), *D, P(dx), *S, 2, None)
    got = host(dpad)
    assert np.array_equal(got[:, :3, :, :5], x)
    assert np.all(got[:, 3:] == 7.0)
    back = empty((2,) + S)
    lib().copyUnPad3DAsync(P(back), *S, P(dpad), *D, 2, None)
    assert np.array_equal(host(back), x)


@pytest.mark.parametrize("n,blocks", [(1, 1), (1000, 3), (123457, 61)])
def test_partial_reductions(n, blocks):
    a, b, c = _rand(n, 100), _rand(n, 101), _rand(n, 102)
    da, db, dc = dev(a), dev(b), dev(c)
    out = empty(blocks)
    L = lib()
    L.partialSumAsync(P(da), P(out), blocks, 512, n, None)
    assert abs(host(out).sum() - a.sum()) <= 1e-12 * np.abs(a).sum()
    L.partialSDotAsync(P(da), P(db), P(out), blocks, 512, n, None)
    assert abs(host(out).sum() - (a * b).sum()) <= 1e-12 * np.abs(a * b).sum()
    for fn, want, fin in [("partialMaxAsync", a.max(), np.max), ("partialMinAsync", a.min(), np.min),
                          ("partialMaxAbsAsync", np.abs(a).max(), np.max)]:
        getattr(L, fn)(P(da), P(out), blocks, 512, n, None)
        assert fin(host(out)) == want
    L.partialMaxDiffAsync(P(da), P(db), P(out), blocks, 512, n, None)
    assert host(out).max() == np.abs(a - b).max()
    L.partialMaxSumAsync(P(da), P(db), P(out), blocks, 512, n, None)
    assert host(out).max() == np.abs(a + b).max()
    L.partialMaxNorm3SqAsync(P(da), P(db), P(dc), P(out), blocks, 512, n, None)
    assert host(out).max() == po.lib().or_maxnorm3sq(a.ctypes.data, b.ctypes.data, c.ctypes.data, n)
    L.partialMaxNorm3SqDiffAsync(P(da), P(db), P(dc), P(dc), P(da), P(db), P(out), blocks, 512, n, None)
    want = po.lib().or_maxnorm3sqdiff(a.ctypes.data, b.ctypes.data, c.ctypes.data, c.ctypes.data,
                                      a.ctypes.data, b.ctypes.data, n)
    assert abs(host(out).max() - want) <= 1e-15 * want


def test_bad_arguments_are_reported_not_fatal():
    L = lib()
    L.hs_clear_error()
    L.exchange6Async(None, None, None, None, None, None, None, None, 0, 1, 1, 0, 0, 0, 1.0, 1.0, 1.0, 1.0, None)
    assert L.hs_last_error() != 0
    L.hs_clear_error()
    assert L.hs_last_error() == 0


def test_thermal_normals_match_oracle_stream():
    n = 5000
    ctr = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    ctr[:, 0] = np.arange(n)
    key = 0xA4093822299F31D0
    out = torch.empty(4 * n, dtype=torch.int32, device="cuda")
    dctr = torch.from_numpy(ctr.view(np.int32).copy()).cuda()
    lib().hs_philox4x32_10(P(out), P(dctr), n, key, None)
    got = host(out).view(np.uint32).reshape(n, 4)
    k2 = [key & 0xFFFFFFFF, key >> 32]
    want = np.stack([po.philox4x32_10(ctr[i], k2) for i in range(n)])
    assert np.array_equal(got, want)
    xi = empty(3 * n)
    lib().hs_thermal_normals_f64(P(xi), n, 1234, 3, 17, 2, None)
    want = po.thermal_normals(1234, 3, 17, 2, n).ravel()
    assert np.abs(host(xi) - want).max() <= 1e-12
    # the FP32 normals of the stage kernels (hardware log2/sqrt/sin/cos): a wrong
    # argument unit of v_sin/v_cos (radians for revolutions) would be off by O(1)
    xf = empty(3 * n, dtype=torch.float32)
    lib().hs_thermal_normals_f32(P(xf), n, 1234, 3, 17, 2, None)
    got = host(xf).astype(np.float64)
    assert np.abs(got - want).max() <= 2e-5
    assert np.abs(got - want).mean() <= 1e-6


def test_mfa_helpers_kappa_ts_cp():
    n = 3000
    rng = np.random.default_rng(50)
    T = rng.uniform(1.0, 1200.0, n)
    T[:3] = 0.0
    ms0 = rng.uniform(0.05, 1.0, n)
    ms0T0 = rng.uniform(0.5, 1.0, n)
    ms0T0[5] = 0.0
    Tc = rng.uniform(0.9, 1.1, n)
    S = rng.choice([0.5, 1.0, 3.5], n)
    nm = rng.uniform(0.5, 2.0, n)
    want = po.kappa(T, Tc, S, nm, ms0, ms0T0, 1.0, 1.0, 1044.0, 1.0)
    dT, dTc, dS, dn, dms0, dms0T0 = (dev(x) for x in (T, Tc, S, nm, ms0, ms0T0))
    k = empty(n)
    lib().kappaAsync(P(k), P(dms0), P(dms0T0), P(dT), P(dTc), P(dS), P(dn), 1.0, 1.0, 1044.0, 1.0, n, None)
    # f/(1-f) amplifies rounding of dB/dx near f = 1 (T -> Tc)
    assert relerr(host(k), want) <= 1e-10
    # Ts: the inverse-Brillouin root; Ridders iterations keep the two within ~eps
    ms = rng.uniform(0.0, 1.0, n)
    ms[7] = 0.0
    ms[8] = 1.0
    want = po.ts(n, ms, None, Tc, S, 1.0, 1.0, 1044.0, 1.0)
    dms = dev(ms)
    t = empty(n)
    lib().tsAsync(P(t), P(dms), None, P(dTc), P(dS), 1.0, 1.0, 1044.0, 1.0, n, None)
    assert relerr(host(t), want) <= 1e-8
    # Cp over both branches (T/Td below and above 0.01)
    Tp = rng.uniform(0.5, 2000.0, n)
    Td = rng.uniform(0.5, 1.5, n)
    Td[11] = 0.0
    want = po.cp(Tp, Td, nm, 470.0)
    dTp, dTd = dev(Tp), dev(Td)
    c = empty(n)
    lib().cpAsync(P(c), P(dTp), P(dTd), P(dn), 470.0, n, None)
    assert relerr(host(c), want) <= TOL


def test_energy_flow_and_uniaxial_anisotropy():
    n = 2500
    m, R = _rand((3, n), 60), _rand((3, n), 61) * 1e12
    Tc, S, nm = _rand(n, 62, 500, 1100), _rand(n, 63, 0, 4), _rand(n, 64, 0, 2)
    S[4] = 0.0
    nm[9] = 0.0
    want = po.energy_flow(m, R, Tc, S, nm, 1.0)
    dm, dR, dTc, dS, dn = dev(m), dev(R), dev(Tc), dev(S), dev(nm)
    w = empty(n)
    lib().energyFlowAsync(P(w), P(dm[0]), P(dm[1]), P(dm[2]), P(dR[0]), P(dR[1]), P(dR[2]), P(dTc), P(dS), P(dn),
                          1.0, n, None)
    assert relerr(host(w), want) <= TOL
    Ku, Ms = _rand(n, 65, 0, 1), _rand(n, 66, 0.2, 1)
    Ms[2] = 0.0
    u = [_rand(n, 67 + c) for c in range(3)]
    u[0][6] = u[1][6] = u[2][6] = 0.0
    want = po.uniaxial_anisotropy(m, Ku, Ms, 2 * 5e5 / 1.7e6, u, (1.0, 0.5, 2.0))
    dKu, dMs = dev(Ku), dev(Ms)
    du = [dev(x) for x in u]
    h = empty((3, n))
    lib().uniaxialAnisotropyAsync(P(h[0]), P(h[1]), P(h[2]), P(dm[0]), P(dm[1]), P(dm[2]), P(dKu), P(dMs),
                                  2 * 5e5 / 1.7e6, P(du[0]), 1.0, P(du[1]), 0.5, P(du[2]), 2.0, None, n)
    assert relerr(host(h), want) <= TOL


@pytest.mark.parametrize("N,pbc,masks", [((3, 5, 70), (0, 0, 0), True), ((4, 9, 130), (1, 1, 1), True),
                                         ((2, 16, 8), (0, 1, 0), False), ((1, 3, 65), (1, 1, 1), True)])
def test_llbar_nonlocal00nc(N, pbc, masks):
    rng = np.random.default_rng(N[1] * 7 + N[2])
    n = N[0] * N[1] * N[2]
    H = rng.standard_normal((3, n)) * 1e5
    msk = rng.uniform(0.2, 1.0, n)
    msk[::13] = 0.0
    lam = [rng.uniform(0.5, 1.5, n) for _ in range(3)] if masks else [None] * 3
    lmul = (2e-19, 3e-19, 5e-19)
    cell = (2e-9, 3e-9, 4e-9)
    want = po.llbar_nonlocal00nc(H, msk, lam, lmul, N, cell, pbc)
    dH, dmsk = dev(H), dev(msk)
    dl = [dev(x) if x is not None else None for x in lam]
    t = empty((3, n))
    lib().llbar_nonlocal00nc_async(P(t[0]), P(t[1]), P(t[2]), P(dH[0]), P(dH[1]), P(dH[2]), P(dmsk),
                                   *[P(x) for x in dl], *lmul, *N, *cell, *pbc, None)
    assert relerr(host(t), want) <= TOL


def test_div_mul_pow_and_wavg():
    n = 4000
    a, b, c = _rand(n, 70), _rand(n, 71, 0.1, 2), _rand(n, 72, 0.0, 3.0)
    b[::17] = 0.0
    c[::19] = 0.0
    want = po.div_mul_pow(a, b, c, 1.7)
    da, db, dc = dev(a), dev(b), dev(c)
    d = empty(n)
    lib().divMulPowAsync(P(d), P(da), P(db), P(dc), 1.7, None, n)
    # single-precision power: the two powf implementations may differ by an ulp of float
    assert relerr(host(d), want) <= 3e-7
    lib().divMulPowAsync(P(d), P(da), None, None, 1.7, None, n)
    assert relerr(host(d), po.div_mul_pow(a, None, None, 1.7)) <= TOL
    x0, x1 = _rand(n, 73, 100, 1000), _rand(n, 74, 100, 1000)
    w0, w1, R = _rand(n, 75, 0, 2), _rand(n, 76, 0, 2), _rand(n, 77, 0, 1)
    w0[5] = w1[5] = 0.0
    want = po.wavg(x0, x1, w0, w1, R, 6e3, 3.8e6, 0.5)
    dx0, dx1, dw0, dw1, dR = dev(x0), dev(x1), dev(w0), dev(w1), dev(R)
    lib().wavgAsync(P(d), P(dx0), P(dx1), P(dw0), P(dw1), P(dR), 6e3, 3.8e6, 0.5, None, n)
    assert relerr(host(d), want) <= TOL
