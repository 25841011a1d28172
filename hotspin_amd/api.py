"""Host-side interface of the MI355X hot path, mirroring the reference's Go
wrappers and its LLB-r module composition.

- `Demag` wraps the demag plan (MaxwellPlan + FFTPlanX of
  src/hotspin-core/modules/maxwell_convolution.go and gpu/fftplanX.go).
- `Solver` wraps the resident LLB-r solver; its keyword arguments follow the
  reference's quantities (msat0T0, lex, B_ext, J, n, Tc, T, alpha = the
  llbar/damping/nonconservative/00/local lambda, gamma_LL, Therm_seed,
  cutoff_dt) in internal (z, y, x) order.

Everything runs through libhotspin_amd.so; host arrays are numpy float64.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import HotspinError, check, lib

EULER, HEUN, RK4, RK12, RK23, AM12, BDF_EULER = 0, 1, 2, 3, 4, 5, 6
METHODS = {"euler": EULER, "heun": HEUN, "rk4": RK4, "rk12": RK12, "rk23": RK23, "am12": AM12,
           "bdf-euler": BDF_EULER}


class LLBParams(C.Structure):
    """hs_llb_params of include/hotspin_amd.h."""
    _fields_ = [
        ("N", C.c_int * 3), ("cell", C.c_double * 3), ("pbc", C.c_int * 3),
        ("msat0T0Mul", C.c_double), ("msat0T0", C.c_void_p),
        ("exchange", C.c_int), ("lexMul", C.c_double), ("lex", C.c_void_p),
        ("demag", C.c_int), ("demag_accuracy", C.c_int),
        ("zeeman", C.c_int), ("Bext", C.c_double * 3),
        ("longfield", C.c_int), ("JMul", C.c_double), ("nMul", C.c_double),
        ("TcMul", C.c_double), ("TMul", C.c_double),
        ("J", C.c_void_p), ("n", C.c_void_p), ("Tc", C.c_void_p), ("T", C.c_void_p),
        ("thermal", C.c_int), ("seed", C.c_uint64), ("stream", C.c_uint32), ("cutoff_dt", C.c_double),
        ("gammaLL", C.c_double), ("torque", C.c_int), ("local00nc", C.c_int),
        ("alphaMul", C.c_double * 3), ("alpha", C.c_void_p * 3),
        ("normalize", C.c_int), ("demag_kernel", C.c_int),
    ]


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_long), ("rejected", C.c_long), ("rhs_evals", C.c_long),
                ("t", C.c_double), ("last_err", C.c_double), ("dt", C.c_double), ("iterations", C.c_long)]


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def make_params(N, cell, pbc=(0, 0, 0), **kw):
    """Build LLBParams from keyword arguments (same names as the fields;
    `alpha_mask` = 3 optional masks).  Returns (params, keepalive)."""
    p = LLBParams()
    N = tuple(int(x) for x in N)
    ncell = N[0] * N[1] * N[2]
    p.N[:] = list(N)
    p.cell[:] = [float(c) for c in cell]
    p.pbc[:] = [int(x) for x in pbc]
    keep = []
    for k, val in kw.items():
        if k in ("msat0T0", "lex", "J", "n", "Tc", "T"):
            a = _f64(val)
            if a is not None:
                a = a.reshape(-1)
                if a.size != ncell:
                    raise ValueError("mask %s has %d cells, mesh has %d" % (k, a.size, ncell))
                keep.append(a)
                setattr(p, k, a.ctypes.data)
        elif k == "alpha_mask":
            for c in range(3):
                if val is not None and val[c] is not None:
                    a = _f64(val[c]).reshape(-1)
                    keep.append(a)
                    p.alpha[c] = a.ctypes.data
        elif k in ("Bext", "alphaMul"):
            getattr(p, k)[:] = [float(x) for x in val]
        else:
            setattr(p, k, val)
    return p, keep


class Solver:
    """Resident LLB-r solve on one GPU (FP32 fields, FP64 state)."""

    def __init__(self, N, cell, pbc=(0, 0, 0), stream=None, **kw):
        self.N = tuple(int(x) for x in N)
        self.ncell = self.N[0] * self.N[1] * self.N[2]
        self.params, self._keep = make_params(N, cell, pbc, **kw)
        L = lib()
        self._h = L.hs_solver_create(C.byref(self.params), stream)
        if not self._h:
            check("hs_solver_create")
            raise HotspinError("hs_solver_create failed")
        self.step_count = 0

    def set_m(self, m):
        m = _f64(m).reshape(3, self.ncell)
        rc = lib().hs_solver_set_m(self._h, m.ctypes.data)
        check("hs_solver_set_m")
        return rc

    def get_m(self):
        out = np.empty((3, self.ncell), np.float64)
        lib().hs_solver_get_m(self._h, out.ctypes.data)
        check("hs_solver_get_m")
        return out

    def rhs(self, step=0, attempt=0, dt=0.0):
        k = np.empty((3, self.ncell), np.float64)
        h = np.empty((3, self.ncell), np.float64)
        lib().hs_solver_rhs(self._h, k.ctypes.data, h.ctypes.data, int(step), int(attempt), float(dt))
        check("hs_solver_rhs")
        return k, h

    def energy(self):
        """[E_zeeman, E_ex, E_demag, E_lf, E_total] of the current state in J
        (micromag/energy, modules/mod_energy.go:25-61)."""
        E = np.empty(5, np.float64)
        lib().hs_solver_energy(self._h, E.ctypes.data)
        check("hs_solver_energy")
        return E

    def run_graph(self, nsteps, dt, step0=None):
        """Fixed-step RK4 from a captured hipGraph (same arithmetic as run("rk4"))."""
        s0 = self.step_count if step0 is None else int(step0)
        rc = lib().hs_solver_run_graph(self._h, int(nsteps), s0, float(dt))
        check("hs_solver_run_graph")
        if rc:
            raise HotspinError("hs_solver_run_graph returned %d" % rc)
        self.step_count = s0 + int(nsteps)

    def set_implicit(self, maxAbsErr=1e-4, maxRelErr=1e-3, maxIterErr=1e-5, maxIter=3):
        """Tolerances of "am12" (m_maxAbsError, m_maxRelError) and "bdf-euler"
        (m_maxIterError, m_maxIterations); the defaults are the reference's."""
        rc = lib().hs_solver_set_implicit(self._h, float(maxAbsErr), float(maxRelErr), float(maxIterErr),
                                          int(maxIter))
        check("hs_solver_set_implicit")
        if rc:
            raise HotspinError("hs_solver_set_implicit returned %d" % rc)

    def set_implicit_T(self, maxAbsErr=1e-4, maxRelErr=1e-3):
        """Te and Tp tolerances of "am12" on a two-temperature solve."""
        rc = lib().hs_solver_set_implicit_T(self._h, float(maxAbsErr), float(maxRelErr))
        check("hs_solver_set_implicit_T")
        if rc:
            raise HotspinError("hs_solver_set_implicit_T returned %d" % rc)

    def run(self, method, nsteps, dt, maxErr=1e-5, mindt=1e-38, maxdt=1e38, step0=None):
        meth = METHODS[method] if isinstance(method, str) else int(method)
        d = C.c_double(dt)
        st = RunStats()
        s0 = self.step_count if step0 is None else int(step0)
        rc = lib().hs_solver_run(self._h, meth, int(nsteps), s0, C.byref(d), float(maxErr), float(mindt),
                                 float(maxdt), C.byref(st))
        check("hs_solver_run")
        if rc:
            raise HotspinError("hs_solver_run returned %d" % rc)
        self.step_count = s0 + st.steps
        return d.value, st

    def close(self):
        if getattr(self, "_h", None):
            lib().hs_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(Solver):
    """nsolve independent solves of one mesh stepped by the same launches
    (per-solve temperature multiplier, thermal seed and applied field)."""

    def __init__(self, N, cell, nsolve, pbc=(0, 0, 0), TMuls=None, seeds=None, Bexts=None, stream=None, **kw):
        self.N = tuple(int(x) for x in N)
        self.ncell = self.N[0] * self.N[1] * self.N[2]
        self.nsolve = int(nsolve)
        self.params, self._keep = make_params(N, cell, pbc, **kw)
        arrT = None if TMuls is None else np.ascontiguousarray(TMuls, dtype=np.float64)
        arrS = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint64)
        arrB = None if Bexts is None else np.ascontiguousarray(Bexts, dtype=np.float64).reshape(-1)
        for a, n in ((arrT, self.nsolve), (arrS, self.nsolve), (arrB, 3 * self.nsolve)):
            if a is not None and a.size != n:
                raise ValueError("per-solve array has %d entries, expected %d" % (a.size, n))
        self._keep += [a for a in (arrT, arrS, arrB) if a is not None]
        self._h = lib().hs_batch_create(C.byref(self.params), self.nsolve,
                                        None if arrT is None else arrT.ctypes.data,
                                        None if arrS is None else arrS.ctypes.data,
                                        None if arrB is None else arrB.ctypes.data, stream)
        if not self._h:
            check("hs_batch_create")
            raise HotspinError("hs_batch_create failed")
        self.step_count = 0

    def set_m(self, m):
        m = _f64(m).reshape(self.nsolve, 3, self.ncell)
        lib().hs_solver_set_m(self._h, m.ctypes.data)
        check("hs_solver_set_m")

    def get_m(self):
        out = np.empty((self.nsolve, 3, self.ncell), np.float64)
        lib().hs_solver_get_m(self._h, out.ctypes.data)
        check("hs_solver_get_m")
        return out

    def rhs(self, step=0, attempt=0, dt=0.0):
        k = np.empty((self.nsolve, 3, self.ncell), np.float64)
        h = np.empty((self.nsolve, 3, self.ncell), np.float64)
        lib().hs_solver_rhs(self._h, k.ctypes.data, h.ctypes.data, int(step), int(attempt), float(dt))
        check("hs_solver_rhs")
        return k, h

    def energy(self):
        """(nsolve, 5): [E_zeeman, E_ex, E_demag, E_lf, E_total] of each solve in J."""
        E = np.empty((self.nsolve, 5), np.float64)
        lib().hs_solver_energy(self._h, E.ctypes.data)
        check("hs_solver_energy")
        return E

    def run_graph(self, nsteps, dt, step0=None):
        """Fixed-step RK4 of every solve from a captured hipGraph."""
        s0 = self.step_count if step0 is None else int(step0)
        rc = lib().hs_solver_run_graph(self._h, int(nsteps), s0, float(dt))
        check("hs_solver_run_graph")
        if rc:
            raise HotspinError("hs_solver_run_graph returned %d" % rc)
        self.step_count = s0 + int(nsteps)


class TmParams(C.Structure):
    """hs_tm_params of include/hotspin_amd.h."""
    _fields_ = [(n, C.c_double) for n in ("gammaE", "Cl", "G", "kappa", "Q0", "sigma_r", "sigma_t", "t0", "xc",
                                          "yc", "maxErrT", "Td", "nD")] + [("spin_heat", C.c_int)]


class TMSolver(Solver):
    """LLB-r coupled to the two-temperature model with a laser source: the
    LLB temperature is Te per cell; (m, Te, Tp) advance together by RK23."""

    def __init__(self, N, cell, pbc=(0, 0, 0), tm=None, T0=300.0, stream=None, **kw):
        self.N = tuple(int(x) for x in N)
        self.ncell = self.N[0] * self.N[1] * self.N[2]
        self.params, self._keep = make_params(N, cell, pbc, **kw)
        self.tm = TmParams()
        for k, v in (tm or {}).items():
            setattr(self.tm, k, int(v) if k == "spin_heat" else float(v))
        self._h = lib().hs_solver_create_tm(C.byref(self.params), C.byref(self.tm), float(T0), stream)
        if not self._h:
            check("hs_solver_create_tm")
            raise HotspinError("hs_solver_create_tm failed")
        self.step_count = 0

    def set_T(self, Te=None, Tp=None):
        te = None if Te is None else _f64(Te).reshape(-1)
        tp = None if Tp is None else _f64(Tp).reshape(-1)
        lib().hs_solver_set_T(self._h, None if te is None else te.ctypes.data, None if tp is None else tp.ctypes.data)
        check("hs_solver_set_T")

    def get_T(self):
        te = np.empty(self.ncell)
        tp = np.empty(self.ncell)
        lib().hs_solver_get_T(self._h, te.ctypes.data, tp.ctypes.data)
        check("hs_solver_get_T")
        return te, tp

    @property
    def time(self):
        return lib().hs_solver_time(self._h)

    @time.setter
    def time(self, t):
        lib().hs_solver_set_time(self._h, float(t))

    def tm_rhs(self, t):
        dte = np.empty(self.ncell)
        dtp = np.empty(self.ncell)
        lib().hs_solver_tm_rhs(self._h, float(t), dte.ctypes.data, dtp.ctypes.data)
        check("hs_solver_tm_rhs")
        return dte, dtp


class Slot(C.Structure):
    """hs_slot of include/hotspin_amd.h."""
    _fields_ = [("step", C.c_uint32), ("attempt", C.c_uint32), ("h", C.c_double), ("t", C.c_double),
                ("err", C.c_double), ("maxErr", C.c_double), ("mindt", C.c_double), ("maxdt", C.c_double),
                ("end", C.c_uint32), ("state", C.c_int32), ("rejected", C.c_uint32), ("accepted", C.c_uint32),
                ("home", C.c_uint32), ("retry", C.c_uint32), ("maxErrT", C.c_double)]


SLOT_IDLE, SLOT_RUNNING, SLOT_DONE, SLOT_FAILED = 0, 1, 2, 3


class SlotBatch:
    """A batch whose nsolve solves are slots filled, stepped (adaptive RK23,
    per-slot dt decided on the device) and emptied independently."""

    def __init__(self, N, cell, capacity, pbc=(0, 0, 0), stream=None, **kw):
        self.N = tuple(int(x) for x in N)
        self.ncell = self.N[0] * self.N[1] * self.N[2]
        self.capacity = int(capacity)
        self.params, self._keep = make_params(N, cell, pbc, **kw)
        self._h = lib().hs_batch_create(C.byref(self.params), self.capacity, None, None, None, stream)
        if not self._h:
            check("hs_batch_create")
            raise HotspinError("hs_batch_create failed")
        self._status = (Slot * self.capacity)()

    def admit(self, slot, m, TMul, seed, nsteps, dt, maxErr=1e-5, mindt=1e-20, maxdt=1e-12, Bext=None,
              stream=0, step0=0):
        m = _f64(m).reshape(3, self.ncell)
        b = None if Bext is None else (C.c_double * 3)(*[float(x) for x in Bext])
        rc = lib().hs_slot_admit(self._h, int(slot), m.ctypes.data, float(TMul), int(seed), int(stream), b,
                                 int(step0), int(nsteps), float(dt), float(maxErr), float(mindt), float(maxdt))
        check("hs_slot_admit")
        if rc:
            raise HotspinError("hs_slot_admit returned %d" % rc)

    def advance(self, iters):
        rc = lib().hs_slot_advance(self._h, int(iters))
        check("hs_slot_advance")
        if rc:
            raise HotspinError("hs_slot_advance returned %d" % rc)

    def ready(self):
        return bool(lib().hs_slot_ready(self._h))

    def status(self):
        lib().hs_slot_status(self._h, self._status)
        check("hs_slot_status")
        return self._status

    def get_m(self, slot):
        out = np.empty((3, self.ncell), np.float64)
        lib().hs_slot_get_m(self._h, int(slot), out.ctypes.data)
        check("hs_slot_get_m")
        return out

    def release(self, slot):
        lib().hs_slot_release(self._h, int(slot))
        check("hs_slot_release")

    def footprint(self):
        return int(lib().hs_solver_footprint(self._h))

    def run_to_end(self, chunk=8):
        """Advance until no slot is running; returns the slot records."""
        while True:
            self.advance(chunk)
            st = self.status()
            if not any(x.state == SLOT_RUNNING for x in st):
                return st

    def close(self):
        if getattr(self, "_h", None):
            lib().hs_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Batches of uneven meshes stepped together: one RK4 step of all of them
    is one hipGraph launch (branches on the batches' streams)."""

    def __init__(self, batches):
        self.batches = list(batches)
        arr = (C.c_void_p * len(self.batches))(*[b._h for b in self.batches])
        self._h = lib().hs_ensemble_create(arr, len(self.batches))
        if not self._h:
            check("hs_ensemble_create")
            raise HotspinError("hs_ensemble_create failed")
        self.step_count = 0

    def run_graph(self, nsteps, dt, step0=None):
        s0 = self.step_count if step0 is None else int(step0)
        rc = lib().hs_ensemble_run_graph(self._h, int(nsteps), s0, float(dt))
        check("hs_ensemble_run_graph")
        if rc:
            raise HotspinError("hs_ensemble_run_graph returned %d" % rc)
        self.step_count = s0 + int(nsteps)

    def close(self):
        if getattr(self, "_h", None):
            lib().hs_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Demag:
    """Demag plan: Kernel_Arne (kernel="arne", the reference's tensor) or the
    Newell tensor (kernel="newell") on the device in FP64 + pruned FFT
    convolution."""

    def __init__(self, N, cell, pbc=(0, 0, 0), accuracy=6, stream=None, kernel="arne"):
        self.N = tuple(int(x) for x in N)
        n = (C.c_int * 3)(*self.N)
        c = (C.c_double * 3)(*[float(x) for x in cell])
        p = (C.c_int * 3)(*[int(x) for x in pbc])
        if kernel == "newell":
            self._h = lib().hs_demag_create_newell(n, c, p, stream)
        elif kernel == "arne":
            self._h = lib().hs_demag_create(n, c, p, int(accuracy), stream)
        else:
            raise ValueError("kernel must be 'arne' or 'newell'")
        if not self._h:
            check("hs_demag_create")
            raise HotspinError("hs_demag_create failed")
        P = (C.c_int * 3)()
        K = (C.c_int * 3)()
        lib().hs_demag_sizes(self._h, P, K)
        self.P = tuple(P)
        self.K = tuple(K)

    def kernel(self):
        out = np.empty((6,) + self.K, np.float64)
        lib().hs_demag_kernel_f64(self._h, out.ctypes.data)
        check("hs_demag_kernel_f64")
        return out

    def max_imag_ratio(self):
        return lib().hs_demag_max_imag_ratio(self._h)

    def field_f64(self, m_ptr, msk_ptr, msat0T0Mul, B_ptr, stream=None):
        rc = lib().hs_demag_field_f64(self._h, m_ptr, msk_ptr, float(msat0T0Mul), B_ptr, stream)
        check("hs_demag_field_f64")
        return rc

    def field_f64_noalloc(self, m_ptr, msk_ptr, msat0T0Mul, B_ptr, stream=None):
        rc = lib().hs_demag_field_f64_noalloc(self._h, m_ptr, msk_ptr, float(msat0T0Mul), B_ptr, stream)
        check("hs_demag_field_f64_noalloc")
        return rc

    def field_f32(self, m_ptr, msk_ptr, msat0T0Mul, B_ptr, stream=None):
        rc = lib().hs_demag_field_f32(self._h, m_ptr, msk_ptr, float(msat0T0Mul), B_ptr, stream)
        check("hs_demag_field_f32")
        return rc

    def close(self):
        if getattr(self, "_h", None):
            lib().hs_demag_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
