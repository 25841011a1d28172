"""Headline benchmark: cell-steps/s of the LLB-r RK step including demag.

Workload (one MI355X): a single LLB-r solve at T = 300 K on a 512x512x16
mesh (user x, y, z) of 2 nm cubes, demag padded to 1024x1024x32, Ms0 =
1.7e6 A/m, Tc = 1044 K, A = 2.1e-11 J/m, lambda = 0.01, J = 1/2, Philox
thermal field with a fixed seed, initial m = seeded random directions per
16^3-cell block with |m| = m_e(300 K), adaptive RK23 (tolerance 1e-5).
value = cells x accepted steps / wall time of the timed steps, with the state
resident in HBM.

Two more legs report batched vs plain per-solve throughput:
- "ensemble": 512 stochastic solves on four mesh shapes stepped together
  (fixed RK4, one graph launch per step) vs the same solves run one at a
  time through the plain solver (a captured graph per solve; eager beside it);
- "serving": the 2000-request Poisson stream served by the
  continuous-batching loop, least work first under a cap on the running
  cells (p50/p99 latency split into queue wait and service time,
  cell-steps/s) vs the plain solver per request.
A "laser" leg runs the laser-heated LLB-r + two-temperature film (5 ps) and
"laser_recipe" the reference's AM12 recipe on the same film.

Every leg reports its error against the FP64 oracle on the same inputs
("error_vs_oracle", "max_abs_dm_vs_oracle", "max_abs_err_vs_oracle"), and
"cpu_baseline" times the oracle on the host cores (and on one thread).
The "legacy_abi" leg times the FP64 legacy entry points and reports the
optional Newell demag plan beside Kernel_Arne.

Prints one JSON line (see the keys below); progress goes to stderr.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       [--ens-steps K] [--serve-requests R] [--serve-policy sjf|fifo]
       [--serve-inflight-cells C] [--no-ensemble] [--no-serving] [--no-laser] [--no-legacy]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from hotspin_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def demag_pass_bytes(N, P):
    """Algorithmic HBM bytes of each FP32 demag pass per launch (each input
    read once, each output written once, pruned extents)."""
    N0, N1, N2 = N
    P0, P1, P2 = P
    H2 = P2 // 2 + 1
    K0, K1 = P0 // 2 + 1, P1 // 2 + 1
    ncell = N0 * N1 * N2
    x1 = 3 * N0 * N1 * H2 * 8
    x2 = 3 * N0 * P1 * H2 * 8
    return {"F1": 3 * ncell * 4 + x1, "F2": x1 + x2, "F3K": 2 * x2 + 6 * K0 * K1 * H2 * 4,
            "I2": x2 + x1, "I1": x1 + 3 * ncell * 4}


def stage_bytes_rk23(ncell, normalize=False):
    """Algorithmic bytes per cell of the four fused RK23 stage kernels (FP32
    m, B, k, e; FP64 y0, y1)."""
    b = [12 + 12 + 24 + 12 + 12 + (24 if normalize else 0),       # m, B, y0 in; k1, m' out (+y0 out)
         12 + 12 + 24 + 12 + 12,                                  # m, B, y0 in; k2, m' out
         12 + 12 + 24 + 12 + 12 + 24 + 12 + 12,                   # m, B, y0, k1, k2 in; y1, e, m' out
         12 + 12 + 12 + (48 if normalize else 0)]                 # m, B, e in (+y1 rw)
    return [x * ncell for x in b]


def ensemble_leg(steps, warmup, plain_steps=10, check_steps=20):
    """512 solves on four shapes: one ensemble graph launch per RK4 step vs
    the plain solver stepping each solve alone.  The plain path is timed as
    a captured graph per solve (hs_solver_run_graph, the same launch savings
    as the batch) and, beside it, eagerly.  Before the timed run the first
    solve of each shape is checked against the FP64 oracle after
    `check_steps` steps taken inside the mixed ensemble."""
    import torch
    from hotspin_amd.api import Batch, Ensemble, Solver
    from oracle import pyoracle as po
    mat = W.MATERIAL_WIDE_T
    cfg = W.ensemble(512, seed=0)
    dt = 1e-15
    batches, counts, cells = [], [], 0
    t0 = time.time()
    for shape in W.ENSEMBLE_SHAPES:
        qs = [c for c in cfg if c["N"] == shape]
        if not qs:
            continue
        kw = W.llbr_params(T=300.0, seed=0, material=mat)
        b = Batch(shape, W.ENSEMBLE_CELL, len(qs), TMuls=[c["T"] for c in qs], seeds=[c["seed"] for c in qs], **kw)
        m0 = np.stack([W.random_domains(shape, 8, W.equilibrium_m(c["T"], mat["Tc"], mat["J"]),
                                        np.random.default_rng(c["seed"])) for c in qs])
        b.set_m(m0)
        batches.append((shape, qs, b, m0))
        counts.append(len(qs))
        cells += len(qs) * int(np.prod(shape))
    e = Ensemble([b for _, _, b, _ in batches])
    log("ensemble: %d solves, %d cells, set-up %.1f s" % (sum(counts), cells, time.time() - t0))
    # FP64 oracle check of one solve per shape, taken from the mixed ensemble
    e.run_graph(check_steps, dt, step0=0)
    errs = {}
    for shape, qs, b, m0 in batches:
        c = qs[0]
        o = po.LLB(shape, W.ENSEMBLE_CELL, **W.llbr_params(T=c["T"], seed=c["seed"], material=mat))
        mo = m0[0].copy()
        for i in range(check_steps):
            o.step("rk4", mo, dt, i)
        errs["x".join(str(n) for n in shape[::-1])] = float(np.abs(b.get_m()[0] - mo).max())
    log("ensemble: max|dm| vs FP64 oracle after %d steps: %s" % (check_steps, errs))
    for _, _, b, m0 in batches:
        b.set_m(m0)
    e.run_graph(warmup, dt, step0=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.run_graph(steps, dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    batched = cells * steps / el
    bad = sum(int(not np.isfinite(b.get_m()).all(axis=(1, 2)).all()) for _, _, b, _ in batches)
    log("ensemble: %d RK4 steps in %.3f s: %.3e cell-steps/s" % (steps, el, batched))
    # plain path: every solve alone, one captured graph per solve (and eager launches)
    plain_time = eager_time = 0.0
    for shape, qs, _, m0 in batches:
        c = qs[0]
        kw = W.llbr_params(T=c["T"], seed=c["seed"], material=mat)
        s = Solver(shape, W.ENSEMBLE_CELL, **kw)
        s.set_m(m0[0])
        s.run_graph(2, dt, step0=0)
        s.run("rk4", 2, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run_graph(plain_steps, dt)
        torch.cuda.synchronize()
        per_step = (time.perf_counter() - t0) / plain_steps
        t0 = time.perf_counter()
        s.run("rk4", plain_steps, dt)
        torch.cuda.synchronize()
        per_eager = (time.perf_counter() - t0) / plain_steps
        plain_time += len(qs) * per_step * steps
        eager_time += len(qs) * per_eager * steps
        log("ensemble: plain %s %.3f ms/step graph, %.3f ms/step eager" % (shape, per_step * 1e3, per_eager * 1e3))
        s.close()
    plain = cells * steps / plain_time
    eager = cells * steps / eager_time
    e.close()
    for _, _, b, _ in batches:
        b.close()
    return {"workload": "512 LLB-r solves, seeds 0..511, meshes {32x32x32, 64x64x4, 128x128x2, 96x48x8} "
                        "3 nm, T~U(100,1000) K, J=7/2, RK4 dt=1e-15", "solves": sum(counts),
            "solves_per_shape": counts, "cells": cells, "steps": steps, "batched_cell_steps_per_s": batched,
            "plain_cell_steps_per_s": plain, "ratio": batched / plain, "plain_eager_cell_steps_per_s": eager,
            "ratio_vs_eager": batched / eager, "nonfinite_solves": bad,
            "max_abs_dm_vs_oracle": errs,
            "oracle_check": "first solve of each shape after %d RK4 steps inside the mixed ensemble vs the FP64 "
                            "oracle (same seed, T, noise counters)" % check_steps,
            "plain_sample": "%d RK4 steps of one solve per shape as a captured graph per solve (eager beside it), "
                            "scaled by the shape counts" % plain_steps}


def serving_leg(nreq, plain_steps=20, check_steps=5, policy="sjf", inflight_cells=1 << 23):
    """Continuous batching of a Poisson request stream vs the plain solver.
    The first request of each shape class is also solved alone for
    `check_steps` adaptive RK23 steps and compared with the FP64 oracle (a
    request's result in the mix equals its solve alone bit for bit)."""
    import torch
    from hotspin_amd.api import SlotBatch, Solver
    from hotspin_amd.serve import Server
    from oracle import pyoracle as po
    reqs = W.requests(n=nreq, rate=200.0, seed=0)
    srv = Server(keep_results=False, policy=policy, inflight_cells=inflight_cells)
    for r in reqs:  # the client sends the initial state with the request
        r["m0"] = srv._initial_state(r)
    srv.run(reqs, log=log)
    summ = srv.summary()
    log("serving: %d requests in %.2f s: %.3e cell-steps/s, p50 %.2f s p99 %.2f s (service p50 %.2f s p99 %.2f s)" % (
        summ["requests"], summ["wall_s"], summ["cell_steps_per_s"], summ["p50_latency_s"], summ["p99_latency_s"],
        summ["p50_service_s"], summ["p99_service_s"]))
    # plain: each request alone through hs_solver_run RK23 (host-side control); one sample request per class
    rate, errs = {}, {}
    for cls, shape in enumerate(W.SERVE_SHAPES):
        rs = [r for r in reqs if r["cls"] == cls]
        if not rs:
            continue
        r = rs[0]
        kw = W.llbr_params(T=r["T"], seed=r["seed"], material=W.MATERIAL_WIDE_T, Bext=r["Bext"])
        s = Solver(shape, W.SERVE_CELL, **kw)
        s.set_m(r["m0"])
        dt, _ = s.run("rk23", 3, 1e-15, maxErr=srv.tol, mindt=srv.mindt, maxdt=srv.maxdt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st = s.run("rk23", plain_steps, dt, maxErr=srv.tol, mindt=srv.mindt, maxdt=srv.maxdt)
        torch.cuda.synchronize()
        rate[cls] = (time.perf_counter() - t0) / max(st.steps, 1)
        s.close()
        # FP64 oracle against the request solved alone in a slot
        b = SlotBatch(shape, W.SERVE_CELL, 1, **srv.kw)
        b.admit(0, r["m0"], r["T"], r["seed"], check_steps, srv.dt0, maxErr=srv.tol, mindt=srv.mindt,
                maxdt=srv.maxdt, Bext=r["Bext"])
        b.run_to_end(chunk=check_steps)
        mg = b.get_m(0)
        b.close()
        o = po.LLB(shape, W.SERVE_CELL, **kw)
        mo, dto = r["m0"].copy(), srv.dt0
        for i in range(check_steps):
            dto, _ = o.step("rk23", mo, dto, i, maxErr=srv.tol, mindt=srv.mindt, maxdt=srv.maxdt)
        errs["x".join(str(n) for n in shape[::-1])] = float(np.abs(mg - mo).max())
    log("serving: max|dm| vs FP64 oracle after %d RK23 steps: %s" % (check_steps, errs))
    srv.close()
    plain_time = sum(r["steps"] * rate[r["cls"]] for r in reqs)
    plain = summ["cell_steps"] / plain_time
    summ.update({"workload": "%d of 2000 Poisson requests at 200 req/s, cells log-uniform 16^3..256x256x4, "
                             "T~U(0,1.2Tc) (>=0.1Tc), |B_ext|<=100 mT, 200-5000 RK23 steps tol 1e-5, J=7/2" % nreq,
                 "plain_cell_steps_per_s": plain, "ratio": summ["cell_steps_per_s"] / plain,
                 "max_abs_dm_vs_oracle": errs,
                 "oracle_check": "first request of each shape class solved alone in a slot for %d adaptive RK23 "
                                 "steps vs the FP64 oracle" % check_steps,
                 "latency_note": "latency = done - arrival, queue = admission - arrival, service = done - "
                                 "admission; the stream offers ~3e11 cell-steps in 10 s, about 12x what the device "
                                 "serves in that time: requests are admitted within a second and then share the "
                                 "device, so service time, not queue wait, makes up the latency",
                 "plain_sample": "%d RK23 steps of one request per shape class through hs_solver_run" % plain_steps})
    return summ


def laser_leg(chunk=50, fine_chunk=5, check_steps=3):
    """Laser-heated LLB-r + two-temperature model, 256x256x8, 5 ps of RK23.
    Te and Tp are read back every `fine_chunk` steps until 1 ps (the pulse
    peaks at 0.3 ps) and every `chunk` steps after it.  Before the timed run
    the GPU's first `check_steps` coupled RK23 steps are compared with the
    FP64 oracle's."""
    import torch
    from hotspin_amd.api import TMSolver
    from oracle import pyoracle as po
    w, kw, tm, m0 = W.laser()
    ncell = int(np.prod(w["N"]))
    t0 = time.time()
    s = TMSolver(w["N"], w["cell"], tm=tm, T0=w["T0"], **kw)
    s.set_m(m0)
    torch.cuda.synchronize()
    setup = time.time() - t0
    # FP64 oracle check of the first steps (each side takes its own adaptive steps)
    o = po.LLB(w["N"], w["cell"], **kw)
    q = po.tm_params(**tm)
    mo, Te, Tp, to, dto = m0.copy(), np.full(ncell, w["T0"]), np.full(ncell, w["T0"]), 0.0, 1e-15
    for i in range(check_steps):
        dto, to, _ = o.step_tm(q, mo, Te, Tp, to, dto, i, maxErr=w["tol"], mindt=1e-20, maxdt=5e-14)
    s.run("rk23", check_steps, 1e-15, maxErr=w["tol"], mindt=1e-20, maxdt=5e-14, step0=0)
    teg, tpg = s.get_T()
    err = {"m": float(np.abs(s.get_m() - mo).max()), "Te_K": float(np.abs(teg - Te).max()),
           "Tp_K": float(np.abs(tpg - Tp).max()), "t_rel": abs(s.time / to - 1)}
    log("laser: vs FP64 oracle after %d steps: %s" % (check_steps, err))
    s.set_m(m0)
    s.set_T(np.full(ncell, w["T0"]), np.full(ncell, w["T0"]))
    s.time = 0.0
    dt, steps, rejected, peak_te, peak_tp = 1e-15, 0, 0, 0.0, 0.0
    t0 = time.perf_counter()
    while s.time < w["span"]:
        n = fine_chunk if s.time < 1e-12 else chunk
        dt, st = s.run("rk23", n, dt, maxErr=w["tol"], mindt=1e-20, maxdt=5e-14, step0=steps)
        steps += st.steps
        rejected += st.rejected
        te, tp = s.get_T()
        peak_te, peak_tp = max(peak_te, float(te.max())), max(peak_tp, float(tp.max()))
    el = time.perf_counter() - t0
    s.close()
    log("laser: %d steps (%d rejected) to t = %.3g s in %.2f s, peak Te %.0f K, peak Tp %.0f K" % (
        steps, rejected, 5e-12, el, peak_te, peak_tp))
    return {"workload": w["name"], "cells": ncell, "steps": steps, "rejected": rejected,
            "cell_steps_per_s": ncell * steps / el, "wall_s": el, "setup_s": setup,
            "peak_Te_K": peak_te, "target_peak_Te_K": w["Tmax"], "peak_Tp_K": peak_tp,
            "max_abs_err_vs_oracle": err,
            "oracle_check": "first %d adaptive coupled RK23 steps of (m, Te, Tp) vs the FP64 oracle" % check_steps,
            "note": "wall time includes reading Te, Tp back every %d steps to 1 ps and every %d after" % (
                fine_chunk, chunk)}


def laser_recipe_leg(nsteps=200):
    """The same film and pump with the couplings and solver of the reference's
    laser recipe (examples/laser-demag.py:30-101): Debye lattice capacity
    (Td = 470 K, nD = the spin density), spin heating Qs = ws - qs into Qe, no
    thermal field (the recipe loads none: with one, qs grows as 1/dt), and
    AM12 of (Te, Tp, m) with absolute and relative bounds 1e-5 for all three.
    The demag field stays on.  A fixed number of steps from t = 0."""
    import torch
    from hotspin_amd.api import TMSolver
    w, kw, tm, m0 = W.laser()
    tm.update(Td=470.0, nD=kw["nMul"], spin_heat=1)
    kw["thermal"] = 0
    ncell = int(np.prod(w["N"]))
    s = TMSolver(w["N"], w["cell"], tm=tm, T0=w["T0"], **kw)
    s.set_m(m0)
    s.set_implicit(maxAbsErr=1e-5, maxRelErr=1e-5)
    s.set_implicit_T(maxAbsErr=1e-5, maxRelErr=1e-5)
    dt, _ = s.run("am12", 5, 1e-16, mindt=1e-20, maxdt=5e-14)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt, st = s.run("am12", nsteps, dt, mindt=1e-20, maxdt=5e-14)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    te, tp = s.get_T()
    t_end = s.time
    s.close()
    log("laser recipe: %d AM12 steps (%d fixed-point iterations) in %.2f s, t = %.3g s, peak Te %.1f K" % (
        st.steps, st.iterations, el, t_end, float(te.max())))
    return {"workload": "laser recipe: LLB-r + 2TM 256x256x8 cells 4x4x2nm, %d AM12 steps from t = 0 (bounds "
                        "1e-5), Debye Cp, spin heating, no thermal field" % nsteps,
            "cells": ncell, "steps": st.steps, "iterations": st.iterations, "rejected": st.rejected,
            "cell_steps_per_s": ncell * st.steps / el, "wall_s": el, "t_end_s": t_end,
            "peak_Te_K": float(te.max()), "peak_Tp_K": float(tp.max())}

def legacy_leg(N=(16, 512, 512), reps=20):
    """The FP64 libhotspin entry points an unchanged Go core calls, one at a
    time on the 512x512x16 mesh (exchange6.cu, long_field.cu, reduce.cu and
    the FP64 demag convolution of MaxwellPlan.UpdateB): time per call from
    HIP events on the null stream (the legacy entry points take a stream
    argument; None is the null stream), and algorithmic GB/s where the bytes
    are simple (each input read once, each output written once)."""
    import ctypes as C

    import torch
    from hotspin_amd._lib import lib
    from hotspin_amd.api import Demag

    L = lib()
    n = N[0] * N[1] * N[2]
    g = torch.Generator(device="cuda").manual_seed(7)
    m = torch.randn((3, n), dtype=torch.float64, device="cuda", generator=g)
    m /= m.norm(dim=0)
    m *= 0.8
    h = torch.empty_like(m)
    Ts = torch.full((n,), 300.0, dtype=torch.float64, device="cuda")
    nb = 1024
    part = torch.empty(nb, dtype=torch.float64, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    res = {}

    def timed(name, fn, nbytes):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[name] = {"us_per_call": ms * 1e3}
        if nbytes:
            res[name]["algorithmic_bytes"] = nbytes
            res[name]["GBps"] = nbytes / (ms * 1e-3) / 1e9
        log("bench: legacy %-22s %8.1f us%s" % (name, ms * 1e3, ("  %5.0f GB/s" % res[name]["GBps"]) if nbytes else ""))

    pre = (1.0 / 4e-18,) * 3
    timed("exchange6Async", lambda: L.exchange6Async(P(h[0]), P(h[1]), P(h[2]), P(m[0]), P(m[1]), P(m[2]), None,
                                                     None, N[0], N[1], N[2], 0, 0, 0, 1.7e6, *pre, None), 48 * n)
    timed("long_field_async", lambda: L.long_field_async(P(h[0]), P(h[1]), P(h[2]), P(m[0]), P(m[1]), P(m[2]), None,
                                                         None, None, None, P(Ts), 1.7e6, 0.5, 1e28, 1044.0, 1.0, n,
                                                         None), 56 * n)
    timed("partialMaxAbsAsync", lambda: L.partialMaxAbsAsync(P(m), P(part), nb, 256, 3 * n, None), 24 * n)
    t0 = time.time()
    d = Demag(N, (2e-9,) * 3)
    torch.cuda.synchronize()
    t_arne = time.time() - t0
    timed("hs_demag_field_f64", lambda: d.field_f64(P(m), None, 1.7e6, P(h)), None)
    # the optional Newell tensor beside Kernel_Arne: plan set-up and the FP64
    # field of the same magnetisation
    t0 = time.time()
    dn = Demag(N, (2e-9,) * 3, kernel="newell")
    torch.cuda.synchronize()
    t_newell = time.time() - t0
    hn = torch.empty_like(h)
    d.field_f64(P(m), None, 1.7e6, P(h))
    dn.field_f64(P(m), None, 1.7e6, P(hn))
    torch.cuda.synchronize()
    diff = ((hn - h).abs().max() / h.abs().max()).item()
    log("bench: Newell plan set-up %.2f s (Kernel_Arne %.2f s), FP64 field max|dB|/max|B| vs Kernel_Arne %.1e"
        % (t_newell, t_arne, diff))
    del d, dn
    return {"workload": "FP64 legacy entry points on the 512x512x16 mesh (2 nm cells), %d calls each" % reps,
            "calls": res,
            "newell": {"setup_s": t_newell, "arne_setup_s": t_arne, "field_max_rel_diff_vs_arne": diff,
                       "note": "plan set-up incl. kernel, FP64 FFT and tile tuning; field of a random m, "
                               "difference = Kernel_Arne's integration error at accuracy 6"}}


TRAFFIC_FILE = os.path.join(HERE, "profiles", "llbr_512x512x16_rk23_hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of a pass from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a plain copy); the stage category
    is the mean of the four RK23 stage roles.  None when not measured."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    table = {}
    for line in open(TRAFFIC_FILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7:
            try:
                table[cells[0]] = float(cells[5]) * 1e6
            except ValueError:
                pass
    if kernel == "STAGE":
        roles = [table.get(r) for r in ("BS_1", "BS_2", "BS_3", "BS_4")]
        return None if None in roles else sum(roles) / 4
    return table.get(kernel)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, kw, m, threads, nsteps=5, single_steps=1):
    """FP64 oracle on the host cores: `nsteps` RK23 steps of the same workload
    on `threads` OpenMP threads, then `single_steps` more on one thread.
    Returns (record, final oracle state after the first nsteps, dt history)."""
    from oracle import pyoracle as po
    po.set_threads(threads)
    t0 = time.time()
    o = po.LLB(w["N"], w["cell"], **kw)
    setup = time.time() - t0
    mo = m.copy()
    dt = 1e-15
    t0 = time.time()
    for i in range(nsteps):
        dt, _ = o.step("rk23", mo, dt, i, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    el = time.time() - t0
    m_end, dt_end = mo.copy(), dt
    po.set_threads(1)
    t1 = time.time()
    for i in range(nsteps, nsteps + single_steps):
        dt, _ = o.step("rk23", mo, dt, i, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    el1 = time.time() - t1
    po.set_threads(threads)
    ncell = int(np.prod(w["N"]))
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    rec = {"value": ncell * nsteps / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
           "sample": "%d RK23 step(s) of the same 512x512x16 workload on the FP64 oracle (%.1f s; "
                     "kernel set-up %.1f s excluded)" % (nsteps, el, setup),
           "single_thread": {"value": ncell * single_steps / el1, "unit": "cell-steps/s", "cores": 1,
                             "sample": "%d further RK23 step(s) on one thread (%.1f s)" % (single_steps, el1)},
           "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": affinity,
           "threads_note": "threads = the CPUs this process may use, capped at 16 (the host share of one GPU)"}
    return rec, m_end, dt_end


def copy_calibration():
    """A plain copy of known size through the library: 2^28 doubles (2 GiB)
    read and 2 GiB written per launch, 5 launches."""
    import torch
    from hotspin_amd import _lib
    L = _lib.lib()
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    D = (1, 1 << 14, 1 << 14)
    for _ in range(5):
        L.copyUnPad3DAsync(b.data_ptr(), D[0], D[1], D[2], a.data_ptr(), D[0], D[1], D[2], 1, None)
    torch.cuda.synchronize()
    _lib.check("copy calibration")
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--ens-steps", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-policy", choices=("fifo", "sjf"), default="sjf")
    ap.add_argument("--serve-inflight-cells", type=int, default=1 << 23)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-laser", action="store_true")
    ap.add_argument("--no-legacy", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 2 GiB (copyUnPad3DAsync with equal extents), "
                         "to scale rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs each solve on a single MI355X "
              "(no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import torch  # noqa: F401  (device initialisation, allocator shared with the library)
    from hotspin_amd import _lib
    from hotspin_amd.api import Solver

    w, kw, m, me = W.single()
    ncell = int(np.prod(w["N"]))
    log("bench: building solver (kernel set-up on the GPU)...")
    t0 = time.time()
    s = Solver(w["N"], w["cell"], **kw)
    s.set_m(m)
    torch.cuda.synchronize()
    log("bench: set-up %.2f s" % (time.time() - t0))
    tiles = (C.c_int * 6)()
    _lib.lib().hs_demag_tiles(_lib.lib().hs_solver_demag(s._h), tiles)
    log("bench: demag tiles (W,NT) rows %s cols %s z %s" % (tuple(tiles[0:2]), tuple(tiles[2:4]), tuple(tiles[4:6])))
    stage_variant = {0: "march", 1: "flat", 2: "pair", 3: "flatz"}.get(_lib.lib().hs_solver_stage_variant(s._h), "?")
    log("bench: stage kernel variant %s" % stage_variant)
    dt = 1e-15
    dt, st = s.run("rk23", args.warmup, dt, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    log("bench: warmup done, dt=%.3e rejected=%d" % (dt, st.rejected))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt, st = s.run("rk23", args.steps, dt, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = ncell * st.steps / el
    log("bench: %d steps in %.3f s, %d rejected, %d demag evals, dt=%.3e" % (st.steps, el, st.rejected,
                                                                            st.rhs_evals, dt))
    # per-kernel durations from HIP events on the solver's stream (separate short run)
    L = _lib.lib()
    ncat = 8
    L.hs_prof_enable(1)
    nprof = 5
    s.run("rk23", nprof, dt, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    tot = (C.c_double * ncat)()
    cnt = (C.c_long * ncat)()
    L.hs_prof_read(tot, cnt, ncat)
    L.hs_prof_enable(0)
    names = ["F1", "F2", "F3K", "I2", "I1", "STAGE"]
    P = (32, 1024, 1024)
    pb = demag_pass_bytes(w["N"], P)
    per = {}
    for i, nme in enumerate(names):
        if cnt[i]:
            avg = tot[i] / cnt[i]
            per[nme] = {"avg_ms": avg, "launches": cnt[i], "total_ms": tot[i]}
            if nme in pb:
                per[nme]["GBps"] = pb[nme] / (avg * 1e-3) / 1e9
    sb = stage_bytes_rk23(ncell)
    if "STAGE" in per:
        per["STAGE"]["GBps"] = (sum(sb) / 4) / (per["STAGE"]["avg_ms"] * 1e-3) / 1e9
    dom = max(per, key=lambda k: per[k]["total_ms"])
    dom_bytes = pb.get(dom, sum(sb) / 4)
    ach = dom_bytes / (per[dom]["avg_ms"] * 1e-3) / 1e9
    for k_, v_ in per.items():
        log("bench: kernel %-5s avg %.4f ms x%d  %.0f GB/s" % (k_, v_["avg_ms"], v_["launches"], v_.get("GBps", 0)))
    out = {
        "metric": "cell-steps/s (LLB RK step incl. demag)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": st.steps,
        "warmup": args.warmup,
        "ms_per_step": el / max(st.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w["name"], "cells": ncell, "method": "rk23", "rejected": st.rejected,
                   "demag_tiles": {"rows": list(tiles[0:2]), "cols": list(tiles[2:4]), "z": list(tiles[4:6])},
                   "stage_kernel": stage_variant,
                   "state": "f64", "fields": "f32", "reductions": "f64", "final_dt": dt},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "algorithmic_bytes": dom_bytes},
        "kernels": per,
    }
    # the same steps with the state handed over through host buffers (hs_solver_set_m /
    # hs_solver_get_m, 100 MB each way over PCIe): reported beside `value`, never as it
    mh = m.copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.set_m(mh)
    _, stp = s.run("rk23", args.steps, dt, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    mh = s.get_m()
    el_p = time.perf_counter() - t0
    out["pcie_inclusive"] = {"value": ncell * stp.steps / el_p, "unit": "cell-steps/s", "steps": stp.steps,
                             "note": "set_m + %d RK23 steps + get_m, host FP64 state copied in and out" % stp.steps}
    log("bench: with host hand-over %.3e cell-steps/s" % (ncell * stp.steps / el_p))
    # the reference's own adaptive solver (solver/rk12, solve_rk12.go) on the same workload:
    # host-driven Heun-Euler with the reference's controller and retries
    s.set_m(m)
    dt12, _ = s.run("rk12", 3, 1e-15, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt12, st12 = s.run("rk12", args.steps, dt12, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12)
    torch.cuda.synchronize()
    el12 = time.perf_counter() - t0
    out["rk12"] = {"value": ncell * st12.steps / el12, "unit": "cell-steps/s", "steps": st12.steps,
                   "rejected": st12.rejected, "ms_per_step": el12 / max(st12.steps, 1) * 1e3, "final_dt": dt12,
                   "note": "reference semantics (solve_rk12.go: 2 RHS per attempt, host-side controller); "
                           "its first-order error estimate holds dt far below RK23's, so compare simulated "
                           "time per second (final_dt), not cell-steps/s"}
    log("bench: rk12 %d steps in %.3f s (%d rejected), dt=%.3e" % (st12.steps, el12, st12.rejected, dt12))
    # energy terms of the final state (micromag/energy): one FP64 pass reading the FP64
    # state (24 B/cell) and the FP32 demag field (12 B/cell) after its own demag evaluation
    s.energy()
    t0 = time.perf_counter()
    ne = 10
    for _ in range(ne):
        E = s.energy()
    e_ms = (time.perf_counter() - t0) / ne * 1e3
    out["energy"] = {"E_J": dict(zip(["zeeman", "exchange", "demag", "longfield", "total"], map(float, E))),
                     "ms_per_call": e_ms, "algorithmic_bytes": 36 * ncell,
                     "note": "wall time per hs_solver_energy call incl. the 40-byte copy back; the demag "
                             "field of the state is cached after the first call"}
    log("bench: energy %.4f ms per call, E_total %.4e J" % (e_ms, E[4]))
    if not args.no_cpu:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
        log("bench: cpu baseline (%d threads, then 1 thread)..." % threads)
        out["cpu_baseline"], m_or, dt_or = cpu_baseline(w, kw, m, threads, args.cpu_steps)
        # the same steps on the GPU from the same FP64 state, against the oracle's result
        s.set_m(m)
        dtg, stg = s.run("rk23", args.cpu_steps, 1e-15, maxErr=w["tol"], mindt=1e-18, maxdt=1e-12, step0=0)
        out["error_vs_oracle"] = {"max_abs_dm": float(np.abs(s.get_m() - m_or).max()), "steps": stg.steps,
                                  "dt_rel": abs(dtg / dt_or - 1), "bar": 1e-5,
                                  "note": "%d adaptive RK23 steps of the 512x512x16 workload from the same FP64 "
                                          "state, GPU (FP32 fields) vs the FP64 oracle" % stg.steps}
        log("bench: vs FP64 oracle after %d steps: max|dm| %.3e" % (stg.steps, out["error_vs_oracle"]["max_abs_dm"]))
    s.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ens_steps, 5)
    if not args.no_serving:
        out["serving"] = serving_leg(args.serve_requests, policy=args.serve_policy,
                                     inflight_cells=args.serve_inflight_cells)
    if not args.no_laser:
        out["laser"] = laser_leg()
        out["laser_recipe"] = laser_recipe_leg()
    if not args.no_legacy:
        out["legacy_abi"] = legacy_leg()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
