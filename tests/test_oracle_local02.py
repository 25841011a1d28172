"""Independent pins of the optional LLB-r damping terms of the oracle
(or_llbar_local02nc, or_llbar_local02c), which restate
src/libhotspin/llbar-local02nc.cu:11-73 and llbar-local02c.cu:11-69:

    local02nc: t_c = mu_c (m . H) m_c
    local02c:  t   = (mu o (m x H)) x m        (mu o v = (mu_x v_x, mu_y v_y, mu_z v_z))

Pinned by closed forms and special cases that a dropped term, a wrong sign
or a swapped cross-product operand would fail:
- isotropic mu: local02nc = mu (m.H) m is parallel to m and zero for H
  perpendicular to m; local02c = mu (|m|^2 H - (m.H) m) (BAC-CAB), so it is
  perpendicular to m, zero for H parallel to m, and t.H >= 0 (it turns m
  towards H);
- anisotropic mu on basis vectors: m = e_i, H = e_j gives
  local02c = mu_k e_k x e_i with e_k = e_i x e_j, and local02nc =
  mu_i delta_ij e_i;
- masks: msat0T0 = 0 gives 0; NULL mu masks mean uniform multipliers;
  mask values multiply the multipliers.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def _call(name, m, H, mu_mul, mu_mask=None, msk=None):
    n = m.shape[1]
    t = np.empty((3, n))
    m = np.ascontiguousarray(m, np.float64)
    H = np.ascontiguousarray(H, np.float64)
    masks = [None, None, None] if mu_mask is None else [np.ascontiguousarray(x, np.float64) for x in mu_mask]
    getattr(po.lib(), name)(*[t[c].ctypes.data for c in range(3)], *[m[c].ctypes.data for c in range(3)],
                            *[H[c].ctypes.data for c in range(3)], None if msk is None else msk.ctypes.data,
                            *[None if x is None else x.ctypes.data for x in masks], *map(float, mu_mul), n)
    return t


@pytest.fixture
def rand():
    rng = np.random.default_rng(3)
    n = 500
    m = rng.standard_normal((3, n)) * rng.uniform(0.1, 1.0, n)
    H = rng.standard_normal((3, n)) * 1e5
    return m, H


def test_isotropic_closed_forms(rand):
    m, H = rand
    mu = 0.37
    mH = (m * H).sum(axis=0)
    m2 = (m * m).sum(axis=0)
    nc = _call("or_llbar_local02nc", m, H, (mu,) * 3)
    c = _call("or_llbar_local02c", m, H, (mu,) * 3)
    scale = np.abs(H).max()
    np.testing.assert_allclose(nc, mu * mH * m, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(c, mu * (m2 * H - mH * m), rtol=0, atol=1e-12 * scale)
    # geometry: local02c is transverse and turns m towards H
    assert np.abs((c * m).sum(axis=0)).max() <= 1e-12 * scale
    assert ((c * H).sum(axis=0) >= -1e-9 * scale * scale).all()
    # H parallel to m: local02c vanishes; H perpendicular to m: local02nc vanishes
    Hpar = 3.0 * m
    assert np.abs(_call("or_llbar_local02c", m, Hpar, (mu,) * 3)).max() <= 1e-14
    Hperp = np.cross(m.T, np.array([0.3, -0.2, 0.9])).T
    assert np.abs(_call("or_llbar_local02nc", m, Hperp, (mu,) * 3)).max() <= 1e-14


def test_anisotropic_basis_vectors():
    mu = (0.11, 0.23, 0.47)
    E = np.eye(3)
    for i in range(3):
        for j in range(3):
            m = E[i][:, None]
            H = E[j][:, None]
            c = _call("or_llbar_local02c", m, H, mu)[:, 0]
            nc = _call("or_llbar_local02nc", m, H, mu)[:, 0]
            ek = np.cross(E[i], E[j])
            k = int(np.argmax(np.abs(ek))) if ek.any() else 0
            want_c = mu[k] * np.cross(ek, E[i])
            want_nc = mu[i] * (i == j) * E[i]
            np.testing.assert_array_equal(c, want_c)
            np.testing.assert_array_equal(nc, want_nc)


def test_masks(rand):
    m, H = rand
    n = m.shape[1]
    rng = np.random.default_rng(4)
    mus = [rng.uniform(0.5, 2.0, n) for _ in range(3)]
    msk = rng.uniform(0.2, 1.0, n)
    msk[::7] = 0.0
    mul = (0.1, 0.2, 0.3)
    for name in ("or_llbar_local02nc", "or_llbar_local02c"):
        full = _call(name, m, H, mul, mus, msk)
        assert (full[:, ::7] == 0).all()
        # per cell, a masked mu equals a uniform multiplier of the same value
        for cidx in (1, 2, 5):
            want = _call(name, m[:, cidx:cidx + 1], H[:, cidx:cidx + 1],
                         [mul[a] * mus[a][cidx] for a in range(3)])
            np.testing.assert_array_equal(full[:, cidx:cidx + 1], want)
        # the msat0T0 mask only switches cells off: its value does not scale t
        ones = _call(name, m, H, mul, mus, np.ones(n))
        keep = msk != 0
        np.testing.assert_array_equal(full[:, keep], ones[:, keep])
