"""Resident FP32-field / FP64-state solver on the GPU vs the FP64 oracle.

Same module composition and inputs on both sides; the thermal noise is the
same Philox stream.  Bars: RHS relative 2e-5 of max|H| / max|k| (FP32
fields); trajectories max|dm| <= 1e-5 (the drop-in target)."""
import math

import numpy as np
import pytest

from hotspin_amd.api import Solver
from hotspin_amd._lib import check
from oracle import pyoracle as po
from tests.gpu_util import relerr, require_gpu

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * math.pi
GAMMA0 = 2.211e5


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()
    yield
    check("solver")


def llb_problem(N, cell, seed=1, thermal=True, pbc=(0, 0, 0), masks=False, T=300.0):
    """LLB-r composition of examples/p2f.py: exchange6, demag, mfa/longfield,
    temperature/brown, llbar torque + local00nc damping."""
    Ms0, Tc, A, lam, J = 1.7e6, 1044.0, 2.1e-11, 0.01, 0.5
    n_spin = Ms0 / (2.0 * 9.27400968202020202e-24 * J)
    lex = math.sqrt(2 * A / (MU0 * Ms0 * Ms0))
    ncell = N[0] * N[1] * N[2]
    rng = np.random.default_rng(seed)
    kw = dict(msat0T0Mul=Ms0, exchange=1, lexMul=lex, demag=1, longfield=1, JMul=J, nMul=n_spin, TcMul=Tc,
              TMul=T, thermal=1 if thermal else 0, seed=1234, stream=0, cutoff_dt=0.0, gammaLL=GAMMA0,
              torque=1, local00nc=1, alphaMul=(lam, lam, lam), zeeman=1, Bext=(0.01, -0.02, 0.05))
    if masks:
        kw["msat0T0"] = 0.5 + 0.5 * rng.random(ncell)
        kw["lex"] = 0.8 + 0.4 * rng.random(ncell)
        kw["T"] = 0.9 + 0.2 * rng.random(ncell)
    me = 0.8
    m = rng.standard_normal((3, ncell))
    m *= me / np.linalg.norm(m, axis=0)
    return kw, m


def llg_problem(N, cell, seed=2):
    """LLG limit of the reference's SP4 recipe (tests/stdproblem4-llbr-modular.py)."""
    Ms0, A, alpha = 8.0e5, 1.3e-11, 0.02
    lex = math.sqrt(2 * A / (MU0 * Ms0 * Ms0))
    kw = dict(msat0T0Mul=Ms0, exchange=1, lexMul=lex, demag=1, zeeman=1, Bext=(0.0, 0.0043, -0.0246),
              gammaLL=GAMMA0 / (1 + alpha ** 2), torque=1, local00nc=1, alphaMul=(alpha,) * 3, normalize=1)
    ncell = N[0] * N[1] * N[2]
    rng = np.random.default_rng(seed)
    m = np.zeros((3, ncell))
    m[2] = 1.0
    m += 0.2 * rng.standard_normal((3, ncell))
    m /= np.linalg.norm(m, axis=0)
    return kw, m


def _normalized(m):
    return m / np.linalg.norm(m, axis=0)


@pytest.mark.parametrize("masks", [False, True])
@pytest.mark.parametrize("pbc", [(0, 0, 0), (0, 1, 1)])
def test_rhs_matches_oracle(masks, pbc):
    N, cell = (4, 12, 80), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, masks=masks, pbc=pbc)
    o = po.LLB(N, cell, pbc, **kw)
    mo, ko = o.rhs(m, step=5, attempt=1, dt=1e-15)
    Ho = o.heff()
    g = Solver(N, cell, pbc, **kw)
    g.set_m(m)
    kg, Hg = g.rhs(step=5, attempt=1, dt=1e-15)
    assert relerr(Hg, Ho) <= 2e-5
    assert relerr(kg, ko) <= 2e-5


@pytest.mark.parametrize("method,steps,dt", [("rk4", 200, 1e-13), ("rk12", 100, 1e-14), ("heun", 100, 5e-14),
                                             ("euler", 100, 1e-14), ("rk23", 100, 1e-14)])
def test_llg_trajectory_matches_oracle(method, steps, dt):
    N, cell = (1, 16, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    kw, m = llg_problem(N, cell)
    o = po.LLB(N, cell, **kw)
    mo = m.copy()
    dto = dt
    for i in range(steps):
        dto, _ = o.step(method, mo, dto, i, maxErr=1e-4, mindt=1e-16, maxdt=1e-12)
    g = Solver(N, cell, **kw)
    g.set_m(m)
    dtg, st = g.run(method, steps, dt, maxErr=1e-4, mindt=1e-16, maxdt=1e-12)
    mg = g.get_m()
    assert np.abs(_normalized(mg) - _normalized(mo)).max() <= 1e-5
    if method in ("rk12", "rk23"):
        assert abs(dtg - dto) <= 1e-4 * dto  # FP32 error estimates move dt continuously


@pytest.mark.parametrize("thermal", [False, True])
def test_llb_rk4_trajectory_matches_oracle(thermal):
    N, cell = (2, 16, 32), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, thermal=thermal)
    o = po.LLB(N, cell, **kw)
    mo = m.copy()
    for i in range(50):
        o.step("rk4", mo, 2e-15, i)
    g = Solver(N, cell, **kw)
    g.set_m(m)
    g.run("rk4", 50, 2e-15)
    mg = g.get_m()
    assert np.abs(mg - mo).max() <= 1e-5


@pytest.mark.parametrize("masks", [False, True])
def test_llb_rk23_adaptive_matches_oracle(masks):
    """Adaptive RK23 on a deterministic LLB-r solve.  The error estimate is a
    difference of stage slopes (cancellation ~ h|k|/err ~ 1e3), so FP32 fields
    move the proposed dt by ~1e-3 relative per step and two free-running
    adaptive runs drift apart in time.  The parity check therefore replays the
    oracle's accepted dt sequence on the GPU (same RK23 arithmetic, max|dm| <=
    1e-5) and checks, step by step, that the GPU's accept/reject decision and
    proposed dt agree with the oracle's from the same state."""
    N, cell = (2, 16, 32), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, thermal=False, masks=masks)
    o = po.LLB(N, cell, **kw)
    g = Solver(N, cell, **kw)
    mo = m.copy()
    g.set_m(m)
    dt = 1e-15
    for i in range(40):
        m_before = mo.copy()
        dt_in = dt
        dt, sto = o.step("rk23", mo, dt, i, maxErr=1e-5, mindt=1e-18, maxdt=1e-12)
        tries = int(sto[1])
        # the dt the oracle accepted: dt_in shrunk by the rejected attempts
        # (replay the same attempt sequence on the GPU from the same state)
        g.set_m(m_before)
        dtg, stg = g.run("rk23", 1, dt_in, maxErr=1e-5, mindt=1e-18, maxdt=1e-12, step0=i)
        assert stg.rejected == tries - 1, (i, stg.rejected, tries)
        assert abs(dtg / dt - 1) <= 2e-3, (i, dtg, dt)  # observed up to 6.4e-4
        assert np.abs(g.get_m() - mo).max() <= 1e-5, i


@pytest.mark.parametrize("variant", ["march", "flat", "pair", "flatz"])
@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1)])
@pytest.mark.parametrize("masks", [True, False])
def test_stage_kernel_variants_match_oracle(variant, pbc, masks, monkeypatch):
    """Each stage-kernel variant (marching planes, one plane per block, two
    cells per thread, one plane per block in XCD plane order) against the
    oracle: RHS and 20 thermal RK4 steps, on a mesh whose rows are not a
    multiple of the tile widths.  With per-cell masks the general kernel form
    runs; without them the lean form (masks and the general longitudinal
    field compiled out)."""
    monkeypatch.setenv("HS_STAGE_KERNEL", variant)
    N, cell = (5, 12, 160), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, masks=masks, pbc=pbc)
    o = po.LLB(N, cell, pbc, **kw)
    mo, ko = o.rhs(m, step=3, attempt=0, dt=1e-15)
    Ho = o.heff()
    g = Solver(N, cell, pbc, **kw)
    from hotspin_amd._lib import lib
    assert lib().hs_solver_stage_variant(g._h) == {"march": 0, "flat": 1, "pair": 2, "flatz": 3}[variant]
    assert lib().hs_solver_stage_lean(g._h) == (0 if masks else 1)
    g.set_m(m)
    kg, Hg = g.rhs(step=3, attempt=0, dt=1e-15)
    assert relerr(Hg, Ho) <= 2e-5
    assert relerr(kg, ko) <= 2e-5
    mo = m.copy()
    for i in range(20):
        o.step("rk4", mo, 2e-15, i)
    g.run("rk4", 20, 2e-15)
    assert np.abs(g.get_m() - mo).max() <= 1e-5


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1)])
def test_stage_kernel_variants_are_bitwise_identical(pbc, monkeypatch):
    """The stage-kernel variant is picked by timing per process; every
    variant does the same arithmetic per cell, so the choice never changes a
    result: 12 thermal RK23 steps give the same bits with each forced."""
    N, cell = (5, 12, 160), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, masks=True, pbc=pbc)
    got = {}
    for variant in ("march", "flat", "pair", "flatz"):
        monkeypatch.setenv("HS_STAGE_KERNEL", variant)
        g = Solver(N, cell, pbc, **kw)
        g.set_m(m)
        g.run("rk23", 12, 1e-15, maxErr=1e-5, mindt=1e-18, maxdt=1e-12)
        got[variant] = g.get_m()
        g.close()
    for variant, mv in got.items():
        assert np.isfinite(mv).all()
        assert np.array_equal(mv, got["flat"]), (variant, np.abs(mv - got["flat"]).max())


@pytest.mark.parametrize("variant", ["march", "flat", "pair", "flatz"])
@pytest.mark.parametrize("pbc", [(0, 0, 0), (0, 1, 1)])
def test_lean_stage_form_equals_general_form(variant, pbc, monkeypatch):
    """The lean stage kernels against the general form (HS_STAGE_LEAN=0) on
    a mask-free thermal LLB-r problem: the RHS and every stepping role
    (Euler, RK12, RK4, RK23) give the same bits."""
    from hotspin_amd._lib import lib
    monkeypatch.setenv("HS_STAGE_KERNEL", variant)
    N, cell = (5, 12, 160), (2e-9, 2e-9, 2e-9)
    kw, m = llb_problem(N, cell, masks=False, pbc=pbc)
    got = {}
    for lean in ("1", "0"):
        monkeypatch.setenv("HS_STAGE_LEAN", lean)
        g = Solver(N, cell, pbc, **kw)
        assert lib().hs_solver_stage_lean(g._h) == int(lean)
        out = []
        g.set_m(m)
        out += list(g.rhs(step=3, attempt=0, dt=1e-15))
        for meth, n in (("euler", 3), ("rk12", 3), ("rk4", 4), ("rk23", 8)):
            g.set_m(m)
            g.run(meth, n, 1e-15, maxErr=1e-5, mindt=1e-18, maxdt=1e-12, step0=0)
            out.append(g.get_m())
        got[lean] = out
        g.close()
    for a, b in zip(got["1"], got["0"]):
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), np.abs(a - b).max()


@pytest.mark.parametrize("variant", ["march", "flat", "pair"])
def test_rk23_error_max_sees_every_lane(variant, monkeypatch):
    """The RK23 error estimate is a max over every cell (wave, block, then
    device reductions).  On a state at the equilibrium magnetisation (no
    demag, no applied field: zero error in the oracle), one cell shortened
    and tilted dominates the error; placed at k = 0, 5, 21, 42, 63
    (different lanes of a wave) and in different rows and planes, the GPU's
    error estimate of one step equals the oracle's."""
    from hotspin_amd import workloads as W
    monkeypatch.setenv("HS_STAGE_KERNEL", variant)
    N, cell = (2, 8, 64), (2e-9, 2e-9, 2e-9)
    kw, _ = llb_problem(N, cell, thermal=False)
    kw.update(demag=0, zeeman=0)
    me = W.equilibrium_m(kw["TMul"], kw["TcMul"], kw["JMul"])
    g = Solver(N, cell, **kw)
    o = po.LLB(N, cell, **kw)
    for (i, j, k) in ((0, 0, 0), (1, 3, 5), (0, 6, 21), (1, 7, 42), (0, 2, 63)):
        m = np.zeros((3, N[0] * N[1] * N[2]))
        m[2] = me
        c = (i * N[1] + j) * N[2] + k
        m[:, c] = (0.9 * math.sin(1.2), 0.0, 0.9 * math.cos(1.2))
        mo = m.copy()
        _, sto = o.step("rk23", mo, 1e-15, 0, maxErr=1.0, mindt=1e-18, maxdt=1e-15)
        g.set_m(m)
        _, stg = g.run("rk23", 1, 1e-15, maxErr=1.0, mindt=1e-18, maxdt=1e-15, step0=0)
        assert sto[0] > 1e-6
        assert abs(stg.last_err / sto[0] - 1) <= 1e-2, ((i, j, k), stg.last_err, sto[0])


@pytest.mark.parametrize("pbc", [(0, 0, 0), (0, 2, 0)])
def test_rhs_and_rk4_with_newell_demag_match_oracle(pbc):
    """The solver with the Newell demag tensor (demag_kernel = 1) against the
    oracle with the same tensor: H_eff and k, then 10 thermal RK4 steps."""
    N, cell = (4, 16, 24), (2e-9, 3e-9, 2.5e-9)
    kw, m = llb_problem(N, cell, pbc=pbc)
    kw["demag_kernel"] = 1
    o = po.LLB(N, cell, pbc, **kw)
    mo, ko = o.rhs(m, step=3, attempt=0, dt=1e-15)
    Ho = o.heff()
    g = Solver(N, cell, pbc, **kw)
    g.set_m(m)
    kg, Hg = g.rhs(step=3, attempt=0, dt=1e-15)
    assert relerr(Hg, Ho) <= 2e-5
    assert relerr(kg, ko) <= 2e-5
    mo = m.copy()
    for i in range(10):
        o.step("rk4", mo, 2e-15, i)
    g.set_m(m)
    g.run("rk4", 10, 2e-15, step0=0)
    assert np.abs(g.get_m() - mo).max() <= 1e-5
