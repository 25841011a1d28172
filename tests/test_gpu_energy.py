"""Energy terms of the resident state (hs_solver_energy) against the FP64
oracle (micromag/energy, modules/mod_energy.go:25-61,98-116).

m, H_ex and H_lf are FP64 on both sides, so E_zeeman, E_ex and E_lf agree to
1e-12, 1e-10 and 1e-9 relative (the sums run in another order; the masks are
rounded to FP32 for both, as the GPU stores them).  B is the FP32 demag of
the state: E_demag to 1e-5.  A batch's terms are those of its solves alone."""
import numpy as np
import pytest

from hotspin_amd.api import Batch, Solver
from hotspin_amd._lib import check
from oracle import pyoracle as po
from tests.gpu_util import require_gpu
from tests.test_gpu_batch import _inputs
from tests.test_gpu_solver import llb_problem, llg_problem

pytestmark = pytest.mark.gpu

TOL = (1e-12, 1e-10, 1e-5, 1e-9)


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("energy")


def _agree(Eg, Eo):
    for t in range(4):
        assert abs(Eg[t] - Eo[t]) <= TOL[t] * abs(Eo[t]), (t, Eg, Eo)
    assert abs(Eg[4] - Eg[:4].sum()) <= 1e-14 * np.abs(Eg[:4]).sum()


def _f32_masks(kw):
    for key in ("msat0T0", "lex", "T"):
        if key in kw:
            kw[key] = np.asarray(kw[key], np.float32).astype(np.float64)


@pytest.mark.parametrize("masks", [False, True])
@pytest.mark.parametrize("pbc", [(0, 0, 0), (0, 1, 1)])
def test_llb_energy_matches_oracle(masks, pbc):
    N, cell = (4, 12, 80), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, masks=masks, pbc=pbc)
    _f32_masks(kw)
    o = po.LLB(N, cell, pbc, **kw)
    mo, _ = o.rhs(m)
    Eo = o.energy(mo)
    assert np.all(Eo[:4] != 0.0)
    g = Solver(N, cell, pbc, **kw)
    g.set_m(m)
    _agree(g.energy(), Eo)


def test_llg_energy_normalises_the_state():
    N, cell = (1, 16, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    kw, m = llg_problem(N, cell)
    m = m * np.linspace(0.7, 1.3, m.shape[1])  # not yet normalised: the LLG limit reads m / |m|
    o = po.LLB(N, cell, **kw)
    mo, _ = o.rhs(m)
    Eo = o.energy(mo)
    g = Solver(N, cell, **kw)
    g.set_m(m)
    _agree(g.energy(), Eo)


def test_energy_after_steps_and_per_solve_in_a_batch():
    N, cell = (2, 16, 32), (2e-9, 2e-9, 2e-9)
    nsolve = 3
    kw, m, T, seeds, B = _inputs(nsolve, N, cell)
    T = np.asarray(T, np.float32).astype(np.float64)  # a batch keeps its per-solve T in FP32
    b = Batch(N, cell, nsolve, TMuls=T, seeds=seeds, Bexts=B, **kw)
    b.set_m(m)
    b.run("rk4", 5, 2e-15)
    Eb = b.energy()
    mb = b.get_m()
    for q in range(nsolve):
        kq = dict(kw)
        kq.update(TMul=float(T[q]), seed=int(seeds[q]), Bext=tuple(B[q]))
        o = po.LLB(N, cell, **kq)
        mo, _ = o.rhs(mb[q])
        _agree(Eb[q], o.energy(mo))
