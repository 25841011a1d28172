"""hs_demag_field_f64_noalloc (FP64 spectra in the library's static device
array, no allocation) against hs_demag_field_f64 (plan-allocated spectra):
the same kernels, so the fields agree bit for bit, and both stay within
1e-12 of the FP64 oracle.  Two calls in a row, the second on a smaller mesh,
so the second reads a scratch array that the first left full of spectra; a
mesh whose spectra exceed the array is refused before any launch."""
import numpy as np
import pytest

from hotspin_amd._lib import HotspinError, check
from hotspin_amd.api import Demag
from oracle import pyoracle as po
from tests.gpu_util import P, dev, empty, host, relerr, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("no_alloc")


def _case(N, cell, pbc, seed):
    rng = np.random.default_rng(seed)
    n = N[0] * N[1] * N[2]
    m = rng.standard_normal((3, n))
    msk = rng.random(n)
    d = Demag(N, cell, pbc)
    dm, dmsk = dev(m), dev(msk)
    Bn = empty((3, n))
    Bn.fill_(np.nan)
    print("no_alloc: N=%s pbc=%s" % (N, pbc), flush=True)
    d.field_f64_noalloc(P(dm), P(dmsk), 8e5, P(Bn))  # before the plan has FP64 spectra of its own
    Bref = empty((3, n))
    d.field_f64(P(dm), P(dmsk), 8e5, P(Bref))
    got, ref = host(Bn), host(Bref)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    want = po.Demag(N, cell, pbc).B(m, msk, 8e5).reshape(3, n)
    assert relerr(got, want) <= 1e-12
    d.close()


def test_noalloc_matches_allocating_entry_two_calls():
    _case((8, 48, 96), (3e-9, 3e-9, 3e-9), (0, 0, 0), 1)
    _case((4, 6, 8), (2e-9, 3e-9, 2.5e-9), (0, 0, 0), 2)
    _case((4, 8, 16), (1e-9, 2e-9, 4e-9), (2, 0, 3), 3)
    _case((7, 8, 10), (2e-9, 2e-9, 2e-9), (0, 0, 0), 4)  # run-time axis-0 length: the P0-row spectrum too


def test_noalloc_refuses_meshes_beyond_the_static_array():
    d = Demag((16, 256, 256), (2e-9, 2e-9, 2e-9))  # 48 * 16 * (256 + 512) * 272 B = 160 MB of spectra
    m = empty((3, 16 * 256 * 256))
    B = empty((3, 16 * 256 * 256))
    with pytest.raises(HotspinError, match="hs error 10002"):
        d.field_f64_noalloc(P(m), None, 8e5, P(B))
    d.close()
