"""Continuous-batching request loop for LLB-r solves on one GPU.

Requests (mesh shape class, T, B_ext, seed, number of RK23 steps) arrive
over time.  Each shape class owns slot batches (api.SlotBatch) on their own
HIP streams; a request is admitted into a free slot of its class, or into a
new batch of its class when the device-memory budget allows one more
(admission by HBM footprint).  The default policy admits the queued request
with the least work (cells x steps) first while the running requests hold
at most `inflight_cells` cells, so short requests are not held behind the
long ones; `policy="fifo"` admits every request on arrival.  Running slots take
adaptive RK23 steps with per-slot dt decided on the device; the loop
enqueues `chunk` iterations per batch, polls the batches' events without
blocking, retires finished slots (reading their state back), and refills
them from the queue, so requests that arrive while others are on the GPU
join the running batches.  The state of every admitted request stays
resident in HBM until it finishes.

The loop is host policy only: every step runs in libhotspin_amd.so.
"""
import collections
import time

import numpy as np

from . import workloads as W
from .api import SLOT_DONE, SLOT_FAILED, SLOT_IDLE, SLOT_RUNNING, SlotBatch


class _Group:
    def __init__(self, cls, shape, capacity, cell, kw):
        self.cls = cls
        self.shape = shape
        self.ncell = shape[0] * shape[1] * shape[2]
        self.batch = SlotBatch(shape, cell, capacity, **kw)
        self.capacity = capacity
        self.req = [None] * capacity  # request in each slot
        self.busy = False
        self.footprint = self.batch.footprint()

    def free_slots(self):
        return [q for q in range(self.capacity) if self.req[q] is None]

    def running(self):
        return any(r is not None for r in self.req)


class Server:
    """shapes: internal (N0, N1, N2) of each class; hbm_budget: bytes the
    slot batches may hold; slots_cells: cells per batch (capacity = slots_cells
    // cells of the class, at least min_slots)."""

    def __init__(self, shapes=W.SERVE_SHAPES, cell=W.SERVE_CELL, material=W.MATERIAL_WIDE_T,
                 hbm_budget=64 << 30, slots_cells=1 << 21, min_slots=4, max_slots=64, chunk=8, tol=1e-5,
                 dt0=1e-15, mindt=1e-20, maxdt=1e-12, keep_results=True, policy="sjf", inflight_cells=1 << 23):
        self.shapes = list(shapes)
        self.cell = cell
        self.material = material
        self.kw = W.llbr_params(T=300.0, seed=0, material=material, Bext=(0.0, 0.0, 0.0))
        self.hbm_budget = int(hbm_budget)
        self.slots_cells = slots_cells
        self.min_slots, self.max_slots = min_slots, max_slots
        self.chunk = chunk
        self.tol, self.dt0, self.mindt, self.maxdt = tol, dt0, mindt, maxdt
        self.keep_results = keep_results
        if policy not in ("fifo", "sjf"):
            raise ValueError("policy must be 'fifo' or 'sjf'")
        # sjf: the queued request with the least work (cells x steps) first,
        # while the running requests hold at most inflight_cells cells (8.4 M,
        # twice what fills the device); fifo: every arrived request is
        # admitted as soon as its class has a slot, and all of them share the
        # device.  On the 2000-request stream of 16^3..256x256x4 solves sjf
        # cut the p50 latency from 66 s to 13 s at the same throughput
        # (cap 4.2 M: 9 s, 1 % less throughput; 16.8 M: 21 s; 33.6 M: 35 s).
        self.policy = policy
        self.inflight_cells = int(inflight_cells)
        self.running_cells = 0
        self.peak_running_cells = 0
        self.groups = []
        self.used = 0
        self.results = {}

    def _capacity(self, cls):
        n = self.shapes[cls][0] * self.shapes[cls][1] * self.shapes[cls][2]
        return int(min(self.max_slots, max(self.min_slots, self.slots_cells // n)))

    def _new_group(self, cls):
        g = _Group(cls, self.shapes[cls], self._capacity(cls), self.cell, self.kw)
        if self.used + g.footprint > self.hbm_budget and any(x.cls == cls for x in self.groups):
            g.batch.close()
            return None
        self.used += g.footprint
        self.groups.append(g)
        return g

    def _initial_state(self, r):
        mat = self.material
        me = max(W.equilibrium_m(r["T"], mat["Tc"], mat["J"]), 0.05)
        n = self.shapes[r["cls"]]
        return W.random_domains(n, 8, me, np.random.default_rng(1_000_000 + r["seed"]))

    def _admit(self, g, q, r):
        m = r["m0"] if r.get("m0") is not None else self._initial_state(r)
        g.batch.admit(q, m, r["T"], r["seed"], r["steps"], self.dt0, maxErr=self.tol, mindt=self.mindt,
                      maxdt=self.maxdt, Bext=r["Bext"])
        g.req[q] = r
        self.running_cells += g.ncell
        self.peak_running_cells = max(self.peak_running_cells, self.running_cells)
        r["t_admit"] = time.perf_counter() - self.t0

    def _retire(self, g, st):
        now = time.perf_counter() - self.t0
        for q in range(g.capacity):
            r = g.req[q]
            if r is None or st[q].state == SLOT_RUNNING:
                continue
            rec = dict(id=r["id"], cls=r["cls"], arrive=r["t"], admit=r["t_admit"], done=now,
                       latency=now - r["t"], steps=int(st[q].accepted), rejected=int(st[q].rejected),
                       failed=st[q].state == SLOT_FAILED, cells=g.ncell, t_sim=float(st[q].t))
            if self.keep_results:
                rec["m"] = g.batch.get_m(q)
            self.results[r["id"]] = rec
            g.batch.release(q)
            g.req[q] = None
            self.running_cells -= g.ncell

    def run(self, requests, time_scale=1.0, log=None):
        """Serve `requests` (workloads.requests()); arrival times are scaled
        by time_scale.  Returns the per-request records."""
        pending = collections.deque(sorted(requests, key=lambda r: r["t"]))
        queues = [collections.deque() for _ in self.shapes]
        self.t0 = time.perf_counter()
        nreq = len(pending)
        last_log = 0.0
        while len(self.results) < nreq:
            now = time.perf_counter() - self.t0
            while pending and pending[0]["t"] * time_scale <= now:
                r = dict(pending.popleft())
                r["t"] *= time_scale
                queues[r["cls"]].append(r)
            progressed = False
            for g in list(self.groups):
                if g.busy:
                    if not g.batch.ready():
                        continue
                    st = g.batch.status()
                    g.busy = False
                    self._retire(g, st)
                    progressed = True
            if self.policy == "sjf":
                progressed |= self._admit_sjf(queues)
            # admission: free slots of existing batches first, then new batches
            for cls, qu in enumerate(queues if self.policy == "fifo" else ()):
                while qu:
                    g = next((x for x in self.groups if x.cls == cls and not x.busy and x.free_slots()), None)
                    if g is None:
                        if any(x.cls == cls and x.free_slots() for x in self.groups):
                            break  # a batch of this class has room once its iterations finish
                        g = self._new_group(cls)
                        if g is None:
                            break
                    for q in g.free_slots():
                        if not qu:
                            break
                        self._admit(g, q, qu.popleft())
                        progressed = True
            for g in self.groups:
                if not g.busy and g.running():
                    g.batch.advance(self.chunk)
                    g.busy = True
                    progressed = True
            if log and now - last_log > 5.0:
                last_log = now
                log("serve: t=%.1fs done %d/%d queued %d running %d" % (
                    now, len(self.results), nreq, sum(len(x) for x in queues),
                    sum(sum(r is not None for r in g.req) for g in self.groups)))
            if not progressed:
                if not any(g.busy for g in self.groups) and pending:
                    time.sleep(max(0.0, min(0.01, pending[0]["t"] * time_scale - now)))
                else:
                    time.sleep(5e-5)
        self.wall = time.perf_counter() - self.t0
        return self.results

    def _slot_for(self, cls):
        """(group, slot) free for a request of class cls, a new group if the
        class has none with room, or None (room only once a busy batch's
        iterations finish, or the HBM budget is spent)."""
        g = next((x for x in self.groups if x.cls == cls and not x.busy and x.free_slots()), None)
        if g is None:
            if any(x.cls == cls and x.free_slots() for x in self.groups):
                return None
            g = self._new_group(cls)
            if g is None:
                return None
        return g, g.free_slots()[0]

    def _admit_sjf(self, queues):
        waiting = sorted((r for qu in queues for r in qu), key=lambda r: self._ncell(r["cls"]) * r["steps"])
        admitted = False
        for r in waiting:
            n = self._ncell(r["cls"])
            if self.running_cells > 0 and self.running_cells + n > self.inflight_cells:
                continue
            gs = self._slot_for(r["cls"])
            if gs is None:
                continue
            queues[r["cls"]].remove(r)
            self._admit(gs[0], gs[1], r)
            admitted = True
        return admitted

    def _ncell(self, cls):
        n = self.shapes[cls]
        return n[0] * n[1] * n[2]

    def summary(self):
        """Throughput and latency percentiles: latency = done - arrival (what a
        client sees), split into the queue wait (admit - arrival) and the
        service time (done - admit, the solve on the device beside the others)."""
        recs = list(self.results.values())
        lat = np.array([r["latency"] for r in recs])
        wait = np.array([r["admit"] - r["arrive"] for r in recs])
        serv = np.array([r["done"] - r["admit"] for r in recs])
        work = sum(r["cells"] * r["steps"] for r in recs)
        return dict(requests=len(recs), failed=int(sum(r["failed"] for r in recs)), wall_s=self.wall,
                    cell_steps=float(work), cell_steps_per_s=work / self.wall,
                    p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                    p50_queue_s=float(np.percentile(wait, 50)), p99_queue_s=float(np.percentile(wait, 99)),
                    p50_service_s=float(np.percentile(serv, 50)), p99_service_s=float(np.percentile(serv, 99)),
                    batches=len(self.groups), hbm_bytes=self.used, policy=self.policy)

    def close(self):
        for g in self.groups:
            g.batch.close()
        self.groups = []
