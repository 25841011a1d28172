"""Batched solves: each solve of a batch is the same alone and in the mix.

A batch steps nsolve solves of one mesh with shared launches (per-solve T,
thermal seed, B_ext).  The per-solve arithmetic is identical to a solve run
alone, so the check is bit-exact; the hipGraph replay of fixed-step RK4 is
bit-exact against the eagerly launched steps; and one solve of the batch is
checked against the FP64 oracle (max|dm| <= 1e-5)."""
import numpy as np
import pytest

from hotspin_amd.api import Batch, Solver
from hotspin_amd._lib import check
from oracle import pyoracle as po
from tests.gpu_util import require_gpu
from hotspin_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("batch")


def _inputs(nsolve, N, cell):
    """Per-solve T in 100..1000 K, |m| = m_e(T); J = 7/2 (stable at dt 2e-15 down to 100 K)."""
    mat = W.MATERIAL_WIDE_T
    kw = W.llbr_params(T=300.0, seed=1234, material=mat, Bext=(0.01, -0.02, 0.05))
    rng = np.random.default_rng(11)
    ncell = N[0] * N[1] * N[2]
    T = rng.uniform(100, 1000, nsolve)
    m = np.stack([W.random_state(ncell, W.equilibrium_m(t, mat["Tc"], mat["J"]), rng) for t in T])
    seeds = np.arange(nsolve, dtype=np.uint64) + 7
    B = rng.uniform(-0.1, 0.1, (nsolve, 3))
    return kw, m, T, seeds, B


@pytest.mark.parametrize("N", [(2, 16, 32), (4, 8, 16)])
def test_batch_equals_solves_alone(N):
    cell = (2e-9, 2e-9, 2e-9)
    nsolve = 5
    kw, m, T, seeds, B = _inputs(nsolve, N, cell)
    b = Batch(N, cell, nsolve, TMuls=T, seeds=seeds, Bexts=B, **kw)
    b.set_m(m)
    b.run("rk4", 6, 2e-15)
    mb = b.get_m()
    assert np.isfinite(mb).all()
    for q in (0, 3):
        kq = dict(kw)
        kq.update(TMul=float(T[q]), seed=int(seeds[q]), Bext=tuple(B[q]))
        s = Solver(N, cell, **kq)
        s.set_m(m[q])
        s.run("rk4", 6, 2e-15)
        assert np.array_equal(s.get_m(), mb[q]), q


def test_graph_replay_is_bit_exact_and_matches_oracle():
    N, cell = (2, 16, 32), (2e-9, 2e-9, 2e-9)
    nsolve = 3
    kw, m, T, seeds, B = _inputs(nsolve, N, cell)
    a = Batch(N, cell, nsolve, TMuls=T, seeds=seeds, Bexts=B, **kw)
    a.set_m(m)
    a.run("rk4", 8, 2e-15)
    g = Batch(N, cell, nsolve, TMuls=T, seeds=seeds, Bexts=B, **kw)
    g.set_m(m)
    g.run_graph(8, 2e-15)
    assert np.isfinite(a.get_m()).all()
    assert np.array_equal(a.get_m(), g.get_m())
    # solve 1 against the FP64 oracle, same noise stream
    ko = dict(kw)
    ko.update(TMul=float(T[1]), seed=int(seeds[1]), Bext=tuple(B[1]))
    o = po.LLB(N, cell, **ko)
    mo = m[1].copy()
    for i in range(8):
        o.step("rk4", mo, 2e-15, i)
    assert np.abs(g.get_m()[1] - mo).max() <= 1e-5
