"""Demag plan on the GPU vs the FP64 oracle and the reference's golden dumps.

- Kernel octant (Kernel_Arne on the device in FP64, own FFT): relative 1e-12.
- FP64 convolution: relative 1e-12 of max|B|.
- FP32 convolution: relative 2e-6 of max|B| (FP32 FFT round-off).
- Golden 64^3 (tests/test-demag/cmp-ref.py): mean over cells of sum_c |dB_c|
  <= 1e-6 T, the reference's own criterion, on the committed samples."""
import os

import numpy as np
import pytest

from hotspin_amd.api import Demag
from hotspin_amd._lib import check
from oracle import pyoracle as po
from tests.golden import regen
from tests.gpu_util import P, dev, empty, host, relerr, require_gpu, torch

pytestmark = pytest.mark.gpu
FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_dumps.npz"))


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("demag")


CASES = [
    ((4, 6, 8), (2e-9, 3e-9, 2.5e-9), (0, 0, 0)),
    ((1, 32, 128), (3e-9, 125e-9 / 32, 500e-9 / 128), (0, 0, 0)),   # SP4-like film
    ((8, 48, 96), (3e-9, 3e-9, 3e-9), (0, 0, 0)),                    # radix-3 sizes
    ((5, 10, 20), (2e-9, 2e-9, 2e-9), (0, 0, 0)),                    # radix-5 sizes
    ((4, 8, 16), (1e-9, 2e-9, 4e-9), (2, 0, 3)),                     # periodic images
    ((3, 16, 10), (1e-9, 1e-9, 1e-9), (0, 1, 0)),
]


@pytest.mark.parametrize("N,cell,pbc", CASES)
def test_kernel_octant_matches_oracle(N, cell, pbc):
    d = Demag(N, cell, pbc)
    o = po.Demag(N, cell, pbc)
    assert d.P == o.P and d.K == o.K
    got = d.kernel()
    want = o.kernel()
    assert relerr(got, want) <= 1e-12
    assert d.max_imag_ratio() < 1e-12


@pytest.mark.parametrize("N,cell,pbc", CASES)
def test_newell_plan_matches_oracle(N, cell, pbc):
    """The Newell tensor plan (hs_demag_create_newell): kernel octant and
    FP64/FP32 fields against the oracle's Newell plan.  The f/g second
    differences cancel, so last-bit differences of asinh/atan between the
    device and the host library are amplified for distant cells (the bar is
    relative to the largest value: up to 2.2e-9 seen, on the 8x48x96 mesh)."""
    d = Demag(N, cell, pbc, kernel="newell")
    o = po.Demag(N, cell, pbc, kernel="newell")
    assert d.P == o.P and d.K == o.K
    assert relerr(d.kernel(), o.kernel()) <= 1e-7
    assert d.max_imag_ratio() < 1e-11  # FP64 FFT round-off of the even kernel (1.3e-12 seen)
    rng = np.random.default_rng(sum(N))
    n = N[0] * N[1] * N[2]
    m = rng.standard_normal((3, n))
    want = o.B(m, None, 8e5).reshape(3, n)
    B = empty((3, n))
    d.field_f64(P(dev(m)), None, 8e5, P(B))
    assert relerr(host(B), want) <= 1e-7
    B32 = empty((3, n), torch.float32)
    d.field_f32(P(dev(m, np.float32)), None, 8e5, P(B32))
    assert relerr(host(B32), want) <= 2e-6


@pytest.mark.parametrize("N,cell,pbc", CASES)
@pytest.mark.parametrize("masked", [False, True])
def test_field_f64_and_f32_match_oracle(N, cell, pbc, masked):
    rng = np.random.default_rng(sum(N) + masked)
    n = N[0] * N[1] * N[2]
    m = rng.standard_normal((3, n))
    msk = rng.random(n) if masked else None
    want = po.Demag(N, cell, pbc).B(m, msk, 8e5).reshape(3, n)
    d = Demag(N, cell, pbc)
    dm = dev(m)
    dmsk = dev(msk) if masked else None
    B = empty((3, n))
    d.field_f64(P(dm), P(dmsk), 8e5, P(B))
    assert relerr(host(B), want) <= 1e-12
    m32 = dev(m, np.float32)
    msk32 = dev(msk, np.float32) if masked else None
    B32 = empty((3, n), torch.float32)
    d.field_f32(P(m32), P(msk32), 8e5, P(B32))
    assert relerr(host(B32), want) <= 2e-6


@pytest.mark.parametrize("pbc,name", [((0, 0, 0), "B_ref"), ((2, 4, 8), "B_ref_pbc")])
def test_golden_demag_fp32(pbc, name):
    """tests/test-demag/cmp-ref.py through the FP32 production path."""
    msk, m = regen.demag_inputs()
    d = Demag((64, 64, 64), regen.CELL_INTERNAL, pbc)
    n = 64 ** 3
    m32 = dev(m.reshape(3, n), np.float32)
    msk32 = dev(msk.reshape(n), np.float32)
    B = empty((3, n), torch.float32)
    d.field_f32(P(m32), P(msk32), 800e3, P(B))
    Bh = host(B).astype(np.float64)
    idx = FIX[name + "_idx"]
    ref = FIX[name + "_val"].astype(np.float64)
    per_cell = np.abs(Bh[:, idx] - ref).sum(axis=0)
    dsum = np.abs(Bh.sum(axis=1) - FIX[name + "_sum"]) / FIX[name + "_sumabs"]
    if pbc == (0, 0, 0):
        assert per_cell.mean() <= 1e-6, per_cell.mean()
        assert dsum.max() <= 2e-6, dsum
    else:
        # the PBC fixture differs from the FP64 oracle by a uniform ~3e-7 T per
        # component (oracle: 9.3e-7 T/cell); FP32 round-off adds on top of it
        assert per_cell.mean() <= 1.2e-6, per_cell.mean()
        assert dsum.max() <= 2e-5, dsum
    # and the FP32 field against the FP64 oracle on the same inputs
    want = po.Demag((64, 64, 64), regen.CELL_INTERNAL, pbc).B(m, msk, 800e3).reshape(3, n)
    assert relerr(Bh, want) <= 2e-6


@pytest.mark.parametrize("N,pbc", [((16, 32, 64), (0, 0, 0)), ((4, 8, 512), (0, 0, 0)), ((16, 32, 64), (0, 2, 0)),
                                   ((2, 512, 16), (0, 0, 0)), ((32, 16, 32), (0, 0, 0)), ((64, 8, 16), (2, 0, 0)),
                                   ((24, 8, 16), (0, 0, 0)), ((20, 8, 16), (0, 0, 0)), ((8, 16, 32), (0, 0, 0)),
                                   ((64, 4, 16), (0, 0, 0))])
def test_every_fp32_pass_variant_matches_fp64(monkeypatch, N, pbc):
    """Each registered FP32 tile variant of the row, column and axis-0 passes
    (forced by index) gives the FP64 field to FP32 accuracy (row length 512
    and column length 1024 are the 512x512x16 workload's; a periodic axis 1
    has no zero padding, so the half-length column variants fall back to the
    full-length pass; axis-0 lengths 16, 40, 48, 64 (from 32 planes, and
    from 64 periodic planes) and 128 also run the passes with two or four
    lanes per pencil)."""
    import ctypes as C
    from hotspin_amd._lib import lib
    cell = (2e-9, 3e-9, 4e-9)
    rng = np.random.default_rng(5)
    ncell = int(np.prod(N))
    m = rng.standard_normal((3, ncell))
    ref = Demag(N, cell, pbc)
    Bd = empty((3, ncell))
    md = dev(m)
    ref.field_f64(P(md), None, 8e5, P(Bd))
    B64 = host(Bd)
    cnt = (C.c_int * 3)()
    lib().hs_demag_variant_count(ref._h, cnt)
    assert min(cnt) >= 1
    mf = dev(m, np.float32)
    for kind in range(3):
        for i in range(cnt[kind]):
            idx = [-1, -1, -1]
            idx[kind] = i
            monkeypatch.setenv("HS_DEMAG_FORCE", ",".join(str(x) for x in idx))
            d = Demag(N, cell, pbc)
            Bf = empty((3, ncell), torch.float32)
            d.field_f32(P(mf), None, 8e5, P(Bf))
            err = np.abs(host(Bf) - B64).max() / np.abs(B64).max()
            assert err <= 2e-6, (kind, i, err)
            d.close()


# Padded lengths without a compiled codelet family run the run-time-length
# passes (fft_any.hpp): prime factors 7, 11 and 13, 25 = 5^2 (a 100-cell axis
# pads to 200 = 8 * 25), an odd x length of a periodic axis (P2 = 9), a one-cell
# x extent (P2 = 1), and such lengths next to compiled ones.
GENERIC = [
    ((11, 8, 8), (2e-9, 2e-9, 2e-9), (0, 0, 0)),       # P0 = 22
    ((7, 7, 14), (2e-9, 3e-9, 2e-9), (0, 0, 0)),       # P = 14, 14, 28 (row FFT of 14)
    ((4, 8, 1), (2e-9, 2e-9, 3e-9), (0, 0, 0)),        # P2 = 1
    ((2, 100, 50), (3e-9, 2e-9, 2e-9), (0, 0, 0)),     # P1 = 200, row FFT of 50
    ((4, 8, 9), (2e-9, 2e-9, 2e-9), (0, 0, 1)),        # periodic x: P2 = 9 (odd)
    ((3, 9, 12), (1e-9, 2e-9, 1e-9), (0, 1, 0)),       # periodic y: P1 = 9
    ((13, 4, 6), (2e-9, 2e-9, 2e-9), (0, 0, 0)),       # P0 = 26
]


@pytest.mark.parametrize("N,cell,pbc", GENERIC)
def test_generic_lengths_match_oracle(N, cell, pbc):
    """Kernel octant (1e-12), FP64 field (1e-12) and FP32 field (2e-6) against
    the oracle, whose FFT takes any length (checked against numpy's)."""
    d = Demag(N, cell, pbc)
    o = po.Demag(N, cell, pbc)
    assert d.P == o.P and d.K == o.K
    assert relerr(d.kernel(), o.kernel()) <= 1e-12
    assert d.max_imag_ratio() < 1e-12
    rng = np.random.default_rng(sum(N))
    n = N[0] * N[1] * N[2]
    m = rng.standard_normal((3, n))
    msk = rng.random(n)
    want = o.B(m, msk, 8e5).reshape(3, n)
    dm, dmsk = dev(m), dev(msk)  # held: a tensor freed after its pointer is taken may be reused
    B = empty((3, n))
    d.field_f64(P(dm), P(dmsk), 8e5, P(B))
    assert relerr(host(B), want) <= 1e-12
    dm32, dmsk32 = dev(m, np.float32), dev(msk, np.float32)
    B32 = empty((3, n), torch.float32)
    d.field_f32(P(dm32), P(dmsk32), 8e5, P(B32))
    assert relerr(host(B32), want) <= 2e-6
    d.close()


def test_generic_length_solves_alone_in_a_batch_and_vs_oracle():
    """An LLB-r batch on a mesh whose axis-0 pass runs at a run-time length
    (P0 = 14, the spectrum of all P0 rows held per solve): each solve equals
    its solve alone bit for bit, and the oracle after 6 RK4 steps."""
    from hotspin_amd import workloads as W
    from hotspin_amd.api import Batch, Solver
    N, cell = (7, 8, 10), (2e-9, 2e-9, 2e-9)
    mat = W.MATERIAL_WIDE_T
    kw = W.llbr_params(T=300.0, seed=5, material=mat, Bext=(0.01, -0.02, 0.03))
    rng = np.random.default_rng(2)
    n = N[0] * N[1] * N[2]
    T = rng.uniform(200, 900, 3)
    m = np.stack([W.random_state(n, W.equilibrium_m(t, mat["Tc"], mat["J"]), rng) for t in T])
    seeds = np.arange(3, dtype=np.uint64) + 11
    b = Batch(N, cell, 3, TMuls=T, seeds=seeds, **kw)
    b.set_m(m)
    b.run("rk4", 6, 2e-15)
    mb = b.get_m()
    kq = dict(kw)
    kq.update(TMul=float(T[1]), seed=int(seeds[1]))
    s = Solver(N, cell, **kq)
    s.set_m(m[1])
    s.run("rk4", 6, 2e-15)
    assert np.array_equal(s.get_m(), mb[1])
    o = po.LLB(N, cell, **kq)
    mo = m[1].copy()
    for i in range(6):
        o.step("rk4", mo, 2e-15, i)
    assert np.abs(mb[1] - mo).max() <= 1e-5


def test_length_beyond_the_generic_passes_is_an_error():
    """Half the padded x length above 4095 (a 4097-cell x extent) is refused
    with HS_ESIZE at plan creation, before anything runs."""
    from hotspin_amd._lib import HotspinError
    with pytest.raises(HotspinError, match="hs error 10002"):  # HS_ESIZE
        Demag((1, 1, 4097), (2e-9, 2e-9, 2e-9))
