"""Headline benchmark: cell-steps/s of the LLB-r RK step including demag.

Workload (one MI355X): a single LLB-r solve at T = 300 K on a 512x512x16
mesh (user x, y, z) of 2 nm cubes, demag padded to 1024x1024x32, Ms0 =
1.7e6 A/m, Tc = 1044 K, A = 2.1e-11 J/m, lambda = 0.01, J = 1/2, Philox
thermal field with a fixed seed, initial m = seeded random directions per
16^3-cell block with |m| = m_e(300 K), adaptive RK23 (tolerance 1e-5).
value = cells x accepted steps / wall time of the timed steps, with the state
resident in HBM.  Prints one JSON line (see the keys below); progress goes to
stderr.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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

MU0 = 4e-7 * math.pi
KB = 1.38064881313131313e-23
MUB = 9.27400968202020202e-24
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(
    name="llb-r T=300K 512x512x16 cells 2nm, demag padded 1024x1024x32, RK23 tol 1e-5",
    N=(16, 512, 512), cell=(2e-9, 2e-9, 2e-9), Ms0=1.7e6, Tc=1044.0, A=2.1e-11, lam=0.01, J=0.5,
    T=300.0, seed=1234, tol=1e-5, block=16,
)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def equilibrium_m(T, Tc, J):
    """m_e(T): root of B_J(3J/(J+1) Tc/T m) = m by bisection (host set-up only)."""
    if T <= 0:
        return 1.0
    if T >= Tc:
        return 0.0
    a, b = (2 * J + 1) / (2 * J), 1 / (2 * J)
    pre = 3 * J / (J + 1) * Tc / T

    def f(m):
        x = pre * m
        return a / math.tanh(a * x) - b / math.tanh(b * x) - m

    lo, hi = 1e-6, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def problem(w):
    N = w["N"]
    Ms0, J = w["Ms0"], w["J"]
    n_spin = Ms0 / (2.0 * MUB * J)
    lex = math.sqrt(2 * w["A"] / (MU0 * Ms0 * Ms0))
    kw = dict(msat0T0Mul=Ms0, exchange=1, lexMul=lex, demag=1, demag_accuracy=6, longfield=1, JMul=J,
              nMul=n_spin, TcMul=w["Tc"], TMul=w["T"], thermal=1, seed=w["seed"], stream=0, cutoff_dt=0.0,
              gammaLL=2.211e5, torque=1, local00nc=1, alphaMul=(w["lam"],) * 3)
    rng = np.random.default_rng(w["seed"])
    bs = w["block"]
    nb = [max(1, n // bs) for n in N]
    dirs = rng.standard_normal((3,) + tuple(nb))
    dirs /= np.linalg.norm(dirs, axis=0)
    me = equilibrium_m(w["T"], w["Tc"], J)
    idx = [np.minimum(np.arange(n) // bs, nb[a] - 1) for a, n in enumerate(N)]
    m = dirs[:, idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]] * me
    return kw, np.ascontiguousarray(m.reshape(3, -1)), me


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


def cpu_baseline(w, kw, m, threads, nsteps=1):
    """FP64 oracle on the host cores: `nsteps` RK23 steps of the same workload."""
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
    ncell = int(np.prod(w["N"]))
    return {"value": ncell * nsteps / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": "%d RK23 step(s) of the same 512x512x16 workload on the FP64 oracle (%.1f s; "
                      "kernel set-up %.1f s excluded)" % (nsteps, el, setup)}


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
    ap.add_argument("--cpu-steps", type=int, default=1)
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

    w = WORKLOAD
    kw, m, me = problem(w)
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
    stage_variant = "flat" if _lib.lib().hs_solver_stage_variant(s._h) == 1 else "march"
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
                     "frac": ach / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": dom_bytes},
        "kernels": per,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log("bench: cpu baseline (%d threads)..." % threads)
        out["cpu_baseline"] = cpu_baseline(w, kw, m, threads, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
