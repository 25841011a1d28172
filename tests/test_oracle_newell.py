"""The oracle's Newell demag tensor (Newell, Williams & Dunlop 1993; an
option beside the reference's Kernel_Arne) against what fixes it
independently: the closed-form self term of a cube (1/3), the trace sum
rule (1 for the self term of any box, 0 between distinct cells: div B = 0),
the parity of each component, the point-dipole limit far away, and
Kernel_Arne (a numerical integration of the same tensor, pinned by the
reference's golden demag dumps) to its integration accuracy."""
import numpy as np
import pytest

from oracle import pyoracle as po


def test_cube_self_term_is_one_third():
    N = po.newell_tensor((0, 0, 0), (2e-9, 2e-9, 2e-9))
    assert np.allclose(N[:3], 1 / 3, rtol=0, atol=1e-15)
    assert np.abs(N[3:]).max() < 1e-15


@pytest.mark.parametrize("h", [(2e-9, 3e-9, 2.5e-9), (1e-9, 1e-9, 4e-9), (3.9e-9, 3.9e-9, 3e-9)])
def test_trace_sum_rule_and_parity(h):
    assert abs(po.newell_tensor((0, 0, 0), h)[:3].sum() - 1) < 1e-14
    rng = np.random.default_rng(3)
    for _ in range(40):
        ijk = rng.integers(-60, 61, 3)
        if not ijk.any():
            continue
        R = ijk * np.asarray(h)
        N = po.newell_tensor(R, h)
        # to the FP64 accuracy of the second differences (up to 1e-5 at the
        # far-field switch, 1e-16 next to the cell)
        assert abs(N[:3].sum()) <= 2e-5 * np.abs(N).max(), (ijk, N)
        if np.abs(ijk).max() <= 3:
            assert abs(N[:3].sum()) <= 1e-12 * np.abs(N).max(), (ijk, N)
        # diagonal even, XY odd in x and y, XZ in x and z, YZ in y and z
        tol = (1e-12 if np.abs(ijk).max() <= 3 else 2e-5) * np.abs(N).max()
        M = po.newell_tensor(R * np.array([-1, 1, 1]), h)
        assert np.abs(M * np.array([1, 1, 1, 1, -1, -1]) - N).max() <= tol, (ijk, M, N)
        M = po.newell_tensor(R * np.array([1, 1, -1]), h)
        assert np.abs(M * np.array([1, 1, 1, -1, -1, 1]) - N).max() <= tol, (ijk, M, N)


@pytest.mark.parametrize("h", [(2e-9, 2e-9, 2e-9), (2e-9, 3e-9, 2.5e-9)])
def test_far_field_is_the_point_dipole(h):
    """N -> -V/(4 pi) (3 R R^T - R^2 I)/R^5 with a relative correction of
    order (h/R)^2: checked both inside the switch radius (second differences)
    and beyond it."""
    V = np.prod(h)
    for n in (12, 24, 60, 200):
        R = np.array([0.48, 0.6, 0.64]) * n * max(h)
        R = np.round(R / np.asarray(h)) * np.asarray(h)
        r2 = R @ R
        T = (3 * np.outer(R, R) - r2 * np.eye(3)) / r2 ** 2.5
        dip = -V / (4 * np.pi) * np.array([T[0, 0], T[1, 1], T[2, 2], T[1, 2], T[0, 2], T[0, 1]])
        N = po.newell_tensor(R, h)
        err = np.abs(N - dip).max() / np.abs(dip).max()
        assert err <= 0.5 * (max(h) ** 2 / r2), (n, err)


@pytest.mark.parametrize("size,cell,pbc", [((16, 16, 16), (2e-9, 3e-9, 2.5e-9), (0, 0, 0)),
                                           ((8, 16, 32), (1e-9, 2e-9, 4e-9), (0, 0, 0)),
                                           ((8, 8, 8), (2e-9, 2e-9, 2e-9), (1, 0, 2))])
def test_newell_agrees_with_kernel_arne(size, cell, pbc):
    """The full padded kernels (wrapped layout, image sums) of the two methods
    agree to Kernel_Arne's integration accuracy at accuracy 6, and the Newell
    kernel is exactly zero where Kernel_Arne's image ranges leave no cell."""
    K = po.kernel_newell(size, cell, pbc)
    A = po.kernel_arne(size, cell, pbc)
    assert np.abs(K - A).max() <= 2e-3 * np.abs(A).max()
    for a in range(3):  # the Nyquist plane of an open even axis holds no displacement
        if pbc[a] == 0 and size[a] % 2 == 0:
            sl = [slice(None)] * 4
            sl[a + 1] = size[a] // 2
            assert not K[tuple(sl)].any() and not A[tuple(sl)].any()
    if not any(pbc):  # self term: -N of a box, trace -1
        assert abs(K[:3, 0, 0, 0].sum() + 1) < 1e-12


def test_demag_plan_with_newell_kernel_matches_uniform_box_limits():
    """A uniformly magnetised cube of 8^3 cells: the mean demag field of the
    Newell plan is -M/3 (cube demag factor), as the sum of all cell-pair
    tensors of a uniformly magnetised body is its shape's demag tensor."""
    N = (8, 8, 8)
    d = po.Demag(N, (2e-9, 2e-9, 2e-9), kernel="newell")
    n = int(np.prod(N))
    m = np.zeros((3, n))
    m[2] = 1.0
    Ms = 8e5
    B = d.B(m, None, Ms).reshape(3, n)
    mu0 = 4e-7 * np.pi
    assert abs(B[2].mean() / (mu0 * Ms) + 1 / 3) < 1e-12
    assert np.abs(B[:2].mean(axis=1)).max() < 1e-12 * mu0 * Ms
