"""ctypes bindings of the FP64 CPU oracle (oracle/libhs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhs_oracle.so")

_lib = None

dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build as _b  # noqa: F401
            _b.build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    i, d, u32, u64, v = C.c_int, C.c_double, C.c_uint32, C.c_uint64, C.c_void_p
    L.or_exchange6.argtypes = [v] * 7 + [i] * 6 + [d] * 4
    L.or_llbar_torque.argtypes = [v] * 10 + [i]
    L.or_llbar_local00nc.argtypes = [v] * 10 + [d] * 3 + [i]
    L.or_llbar_local02nc.argtypes = [v] * 13 + [d] * 3 + [i]
    L.or_llbar_local02c.argtypes = [v] * 13 + [d] * 3 + [i]
    L.or_long_field.argtypes = [v] * 11 + [d] * 5 + [i]
    L.or_temperature_scale_noise.argtypes = [v] * 8 + [d] * 4 + [i]
    L.or_brillouin.argtypes = [v] * 5 + [d] * 4 + [i]
    L.or_normalize.argtypes = [v] * 3 + [i]
    L.or_qspat.argtypes = [v] * 3 + [d] + [i] * 3 + [d] * 3 + [i] * 3
    L.or_qinter.argtypes = [v] * 4 + [d, i]
    L.or_tens_symm_vec_mul.argtypes = [v] * 12 + [d] + [i] * 3
    L.or_kappa.argtypes = [v] * 7 + [d] * 4 + [i]
    L.or_ts.argtypes = [v] * 5 + [d] * 4 + [i]
    L.or_cp.argtypes = [v] * 4 + [d, i]
    L.or_energy_flow.argtypes = [v] * 10 + [d, i]
    L.or_uniaxial_anisotropy.argtypes = [v] * 8 + [d, v, d, v, d, v, d, i]
    L.or_llbar_nonlocal00nc.argtypes = [v] * 5 + [i] * 3 + [d] * 3 + [i] * 3
    L.or_div_mul_pow.argtypes = [v] * 4 + [d, i]
    L.or_wavg.argtypes = [v] * 6 + [d] * 3 + [i]
    for name in ("or_sum", "or_max", "or_min", "or_maxabs"):
        getattr(L, name).argtypes = [v, C.c_long]
        getattr(L, name).restype = d
    for name in ("or_maxdiff", "or_maxsum", "or_sdot"):
        getattr(L, name).argtypes = [v, v, C.c_long]
        getattr(L, name).restype = d
    L.or_maxnorm3sq.argtypes = [v] * 3 + [C.c_long]
    L.or_maxnorm3sq.restype = d
    L.or_maxnorm3sqdiff.argtypes = [v] * 6 + [C.c_long]
    L.or_maxnorm3sqdiff.restype = d
    L.or_fft1d.argtypes = [v, i, i]
    L.or_fft1d.restype = i
    L.or_kernel_arne.argtypes = [v, v, v, i, v]
    L.or_demag_new.argtypes = [v, v, v, i]
    L.or_demag_new.restype = v
    L.or_demag_new_kernel.argtypes = [v, v, v, i, i]
    L.or_demag_new_kernel.restype = v
    L.or_kernel_newell.argtypes = [v, v, v, v]
    L.or_newell_tensor.argtypes = [v, v, v]
    L.or_demag_free.argtypes = [v]
    L.or_demag_sizes.argtypes = [v, v, v]
    L.or_demag_kernel.argtypes = [v]
    L.or_demag_kernel.restype = dp
    L.or_demag_max_imag_ratio.argtypes = [v]
    L.or_demag_max_imag_ratio.restype = d
    L.or_demag_B.argtypes = [v, v, v, d, v]
    L.or_philox4x32_10.argtypes = [v, v, v]
    L.or_thermal_normals.argtypes = [u64, u32, u32, u32, C.c_long, v]
    L.or_llb_new.argtypes = [v]
    L.or_llb_new.restype = v
    L.or_llb_free.argtypes = [v]
    L.or_llb_rhs.argtypes = [v, v, v, u32, u32, d]
    L.or_debug_set_thermal_prefactor_scale.argtypes = [d]
    L.or_tm_rhs.argtypes = [v, v, v, v, v, v, d, v, v, v, v]
    L.or_llb_spin_heat.argtypes = [v, v, v]
    L.or_llb_energy.argtypes = [v, v, v]
    L.or_llbtm_step_rk23.argtypes = [v, v, v, v, v, v, v, d, d, d, u32, v]
    L.or_llbtm_step_rk23.restype = i
    L.or_llbtm_step_am12.argtypes = [v, v, v, v, v, v, v, v, v, d, d, u32, v]
    L.or_llbtm_step_am12.restype = i
    L.or_llb_heff.argtypes = [v]
    L.or_llb_heff.restype = dp
    L.or_llb_bdemag.argtypes = [v]
    L.or_llb_bdemag.restype = dp
    for name in ("or_llb_step_euler", "or_llb_step_heun", "or_llb_step_rk4"):
        getattr(L, name).argtypes = [v, v, d, u32]
        getattr(L, name).restype = i
    for name in ("or_llb_step_rk12", "or_llb_step_rk23"):
        getattr(L, name).argtypes = [v, v, v, d, d, d, u32, v]
        getattr(L, name).restype = i
    L.or_llb_step_am12.argtypes = [v, v, v, d, d, d, d, u32, v]
    L.or_llb_step_am12.restype = i
    L.or_llb_step_bdf_euler.argtypes = [v, v, d, u32, d, i, v]
    L.or_llb_step_bdf_euler.restype = i
    L.or_energy_term.argtypes = [v, v, v, v, v, d, d, C.c_long]
    L.or_energy_term.restype = d
    L.or_set_threads.argtypes = [i]
    L.or_get_threads.restype = i


def ptr(a):
    """Data pointer of a C-contiguous float64 array, or None."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().or_set_threads(int(n))


def set_thermal_prefactor_scale(f):
    """Test hook: multiply the Brown prefactor of the LLB RHS by f (1 = the
    reference's formula); used only to plant a wrong amplitude on purpose."""
    lib().or_debug_set_thermal_prefactor_scale(float(f))


def get_threads():
    return lib().or_get_threads()


def kernel_arne(size, cell, pbc, accuracy=6):
    size = np.asarray(size, dtype=np.int32)
    out = np.zeros((6,) + tuple(int(s) for s in size), dtype=np.float64)
    cs = np.asarray(cell, dtype=np.float64)
    pb = np.asarray(pbc, dtype=np.int32)
    lib().or_kernel_arne(size.ctypes.data, cs.ctypes.data, pb.ctypes.data, int(accuracy), out.ctypes.data)
    return out


def kernel_newell(size, cell, pbc):
    """Full padded Newell kernel -N (6 planes), Kernel_Arne's layout."""
    size = np.asarray(size, dtype=np.int32)
    out = np.zeros((6,) + tuple(int(s) for s in size), dtype=np.float64)
    cs = np.asarray(cell, dtype=np.float64)
    pb = np.asarray(pbc, dtype=np.int32)
    lib().or_kernel_newell(size.ctypes.data, cs.ctypes.data, pb.ctypes.data, out.ctypes.data)
    return out


def newell_tensor(R, h):
    """Newell N (XX,YY,ZZ,YZ,XZ,XY) of displacement R between cells h."""
    r = np.asarray(R, dtype=np.float64)
    hh = np.asarray(h, dtype=np.float64)
    out = np.zeros(6)
    lib().or_newell_tensor(r.ctypes.data, hh.ctypes.data, out.ctypes.data)
    return out


class Demag:
    """FP64 restatement of MaxwellPlan (demag only); kernel "arne" (the
    reference's Kernel_Arne) or "newell"."""

    def __init__(self, N, cell, pbc=(0, 0, 0), accuracy=6, kernel="arne"):
        self.N = tuple(int(n) for n in N)
        n = np.asarray(N, dtype=np.int32)
        cs = np.asarray(cell, dtype=np.float64)
        pb = np.asarray(pbc, dtype=np.int32)
        kind = {"arne": 0, "newell": 1}[kernel]
        self._h = lib().or_demag_new_kernel(n.ctypes.data, cs.ctypes.data, pb.ctypes.data, int(accuracy), kind)
        P = np.zeros(3, np.int32)
        K = np.zeros(3, np.int32)
        lib().or_demag_sizes(self._h, P.ctypes.data, K.ctypes.data)
        self.P = tuple(int(x) for x in P)
        self.K = tuple(int(x) for x in K)

    def kernel(self):
        n = 6 * self.K[0] * self.K[1] * self.K[2]
        p = lib().or_demag_kernel(self._h)
        return np.ctypeslib.as_array(p, shape=(n,)).reshape((6,) + self.K).copy()

    def max_imag_ratio(self):
        return lib().or_demag_max_imag_ratio(self._h)

    def B(self, m, msat0T0_mask, msat0T0Mul):
        m = f64(m).reshape(3, -1)
        mask = f64(msat0T0_mask)
        out = np.empty_like(m)
        lib().or_demag_B(self._h, ptr(m), ptr(mask), float(msat0T0Mul), ptr(out))
        return out.reshape((3,) + self.N)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_demag_free(self._h)
            self._h = None


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def thermal_normals(seed, stream, step, attempt, N):
    out = np.empty(3 * N, np.float64)
    lib().or_thermal_normals(int(seed), int(stream), int(step), int(attempt), int(N), out.ctypes.data)
    return out.reshape(3, N)


class LLBParams(C.Structure):
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


class TmParams(C.Structure):
    """or_tm_params of oracle/hs_oracle.h (two-temperature model + laser)."""
    _fields_ = [(n, C.c_double) for n in ("gammaE", "Cl", "G", "kappa", "Q0", "sigma_r", "sigma_t", "t0", "xc",
                                          "yc", "maxErrT", "Td", "nD")] + [("spin_heat", C.c_int)]


def tm_params(**kw):
    q = TmParams()
    for k, v in kw.items():
        setattr(q, k, int(v) if k == "spin_heat" else float(v))
    return q


def tm_rhs(q, N, cell, pbc, Te, Tp, t, Qs=None):
    """dTe/dt, dTp/dt of the two-temperature model at time t (Qs: the spin term of Qe per cell)."""
    Te = f64(Te).reshape(-1)
    Tp = f64(Tp).reshape(-1)
    n = Te.size
    dTe, dTp, work = np.empty(n), np.empty(n), np.empty(n)
    Na = (C.c_int * 3)(*N)
    ca = (C.c_double * 3)(*cell)
    pa = (C.c_int * 3)(*pbc)
    lib().or_tm_rhs(C.byref(q), Na, ca, pa, ptr(Te), ptr(Tp), float(t), ptr(f64(Qs)), ptr(dTe), ptr(dTp), ptr(work))
    return dTe, dTp


class LLB:
    """FP64 LLB-r problem: RHS assembly and the drivers (euler/heun/rk4/rk12/rk23)."""

    _MASKS = ("msat0T0", "lex", "J", "n", "Tc", "T")

    def __init__(self, N, cell, pbc=(0, 0, 0), **kw):
        p = LLBParams()
        self.N = tuple(int(x) for x in N)
        self.ncell = self.N[0] * self.N[1] * self.N[2]
        p.N[:] = list(self.N)
        p.cell[:] = [float(c) for c in cell]
        p.pbc[:] = [int(x) for x in pbc]
        self._keep = []
        for k, val in kw.items():
            if k in self._MASKS:
                a = f64(val)
                if a is not None:
                    a = a.reshape(-1)
                    assert a.size == self.ncell
                    self._keep.append(a)
                    setattr(p, k, a.ctypes.data)
                continue
            if k == "alpha_mask":
                for c in range(3):
                    if val is not None and val[c] is not None:
                        a = f64(val[c]).reshape(-1)
                        self._keep.append(a)
                        p.alpha[c] = a.ctypes.data
                continue
            if k in ("Bext", "alphaMul"):
                getattr(p, k)[:] = [float(x) for x in val]
                continue
            setattr(p, k, val)
        self.params = p
        self._h = lib().or_llb_new(C.byref(p))

    def rhs(self, m, step=0, attempt=0, dt=0.0):
        m = f64(m).reshape(3, self.ncell).copy()
        k = np.empty_like(m)
        lib().or_llb_rhs(self._h, ptr(m), ptr(k), int(step), int(attempt), float(dt))
        return m, k

    def energy(self, m):
        """[E_zeeman, E_ex, E_demag, E_lf, E_total] at the state of the last rhs(m)
        call (pass the m that rhs returned)."""
        m = f64(m).reshape(3, self.ncell)
        out = np.empty(5)
        lib().or_llb_energy(self._h, ptr(m), ptr(out))
        return out

    def spin_heat(self, m):
        """The spin term of Qe (ws - qs) per cell at the state of the last rhs(m) call
        (pass the m that rhs returned)."""
        m = f64(m).reshape(3, self.ncell)
        out = np.empty(self.ncell)
        lib().or_llb_spin_heat(self._h, ptr(m), ptr(out))
        return out

    def heff(self):
        p = lib().or_llb_heff(self._h)
        return np.ctypeslib.as_array(p, shape=(3 * self.ncell,)).reshape(3, self.ncell).copy()

    def bdemag(self):
        p = lib().or_llb_bdemag(self._h)
        return np.ctypeslib.as_array(p, shape=(3 * self.ncell,)).reshape(3, self.ncell).copy()

    def step(self, method, m, dt, step=0, maxErr=1e-5, mindt=1e-38, maxdt=1e38):
        """Advance m (float64 (3,ncell), modified in place).  Returns (dt_next, stats)."""
        assert m.dtype == np.float64 and m.flags["C_CONTIGUOUS"]
        L = lib()
        if method in ("euler", "heun", "rk4"):
            rc = getattr(L, "or_llb_step_" + method)(self._h, ptr(m), float(dt), int(step))
            if rc < 0:
                raise RuntimeError("oracle step failed")
            return dt, None
        dtb = np.array([dt], np.float64)
        st = np.zeros(2, np.float64)
        rc = getattr(L, "or_llb_step_" + method)(self._h, ptr(m), dtb.ctypes.data, float(maxErr),
                                                 float(mindt), float(maxdt), int(step), st.ctypes.data)
        if rc < 0:
            raise RuntimeError("oracle adaptive step did not converge")
        return float(dtb[0]), st

    def step_am12(self, m, dt, step=0, absErr=1e-4, relErr=1e-3, mindt=1e-38, maxdt=1e38):
        """One adaptive AM12 step (m modified in place).  Returns (dt_next,
        stats[err, attempts, iterations, rhs_evals, dt_used])."""
        assert m.dtype == np.float64 and m.flags["C_CONTIGUOUS"]
        dtb = np.array([dt], np.float64)
        st = np.zeros(5, np.float64)
        rc = lib().or_llb_step_am12(self._h, ptr(m), dtb.ctypes.data, float(absErr), float(relErr), float(mindt),
                                    float(maxdt), int(step), st.ctypes.data)
        if rc < 0:
            raise RuntimeError("oracle AM12 step did not converge")
        return float(dtb[0]), st

    def step_bdf_euler(self, m, dt, step=0, maxIterErr=1e-5, maxIter=3):
        """One backward-Euler step (m modified in place).  Returns
        stats[iterator err, iterations, rhs_evals]."""
        assert m.dtype == np.float64 and m.flags["C_CONTIGUOUS"]
        st = np.zeros(3, np.float64)
        rc = lib().or_llb_step_bdf_euler(self._h, ptr(m), float(dt), int(step), float(maxIterErr), int(maxIter),
                                         st.ctypes.data)
        if rc < 0:
            raise RuntimeError("oracle backward-Euler iterator did not converge")
        return st

    def step_tm_am12(self, q, m, Te, Tp, t, dt, step=0, absErr=(1e-4, 1e-4, 1e-4), relErr=(1e-3, 1e-3, 1e-3),
                     mindt=1e-38, maxdt=1e38):
        """One AM12 step of (Te, Tp, m) with T = Te (arrays modified in place);
        tolerances per equation (Te, Tp, m).  Returns (dt_next, t_next,
        stats[err_Te, err_Tp, err_m, attempts, iterations, dt_used])."""
        for a in (m, Te, Tp):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        dtb = np.array([dt], np.float64)
        tb = np.array([t], np.float64)
        ab = np.array(absErr, np.float64)
        rl = np.array(relErr, np.float64)
        st = np.zeros(6, np.float64)
        rc = lib().or_llbtm_step_am12(self._h, C.byref(q), ptr(m), ptr(Te), ptr(Tp), tb.ctypes.data, dtb.ctypes.data,
                                      ab.ctypes.data, rl.ctypes.data, float(mindt), float(maxdt), int(step),
                                      st.ctypes.data)
        if rc < 0:
            raise RuntimeError("oracle AM12 step did not converge")
        return float(dtb[0]), float(tb[0]), st

    def step_tm(self, q, m, Te, Tp, t, dt, step=0, maxErr=1e-5, mindt=1e-38, maxdt=1e38):
        """One adaptive RK23 step of (m, Te, Tp) with T = Te (arrays modified in
        place).  Returns (dt_next, t_next, stats[err_m, err_Te, err_Tp, attempts])."""
        for a in (m, Te, Tp):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        dtb = np.array([dt], np.float64)
        tb = np.array([t], np.float64)
        st = np.zeros(4, np.float64)
        rc = lib().or_llbtm_step_rk23(self._h, C.byref(q), ptr(m), ptr(Te), ptr(Tp), tb.ctypes.data,
                                      dtb.ctypes.data, float(maxErr), float(mindt), float(maxdt), int(step),
                                      st.ctypes.data)
        if rc < 0:
            raise RuntimeError("oracle adaptive step did not converge")
        return float(dtb[0]), float(tb[0]), st

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_llb_free(self._h)
            self._h = None


def exchange6(m, lex, N, pbc, msat0T0Mul, pre):
    m = f64(m).reshape(3, -1)
    h = np.empty_like(m)
    lx = f64(lex)
    lib().or_exchange6(ptr(h[0]), ptr(h[1]), ptr(h[2]), ptr(m[0]), ptr(m[1]), ptr(m[2]), ptr(lx),
                       int(N[0]), int(N[1]), int(N[2]), int(pbc[0]), int(pbc[1]), int(pbc[2]),
                       float(msat0T0Mul), float(pre[0]), float(pre[1]), float(pre[2]))
    return h


def _planes(a):
    """(pointer array of the 3 planes, the planes) of a (3, n) float64 array."""
    a = f64(a).reshape(3, -1)
    return (C.c_void_p * 3)(*[p.ctypes.data for p in a]), a


def kappa(T, Tc, S, n, msat0, msat0T0, msat0Mul, msat0T0Mul, TcMul, SMul):
    """kappaAsync (kappa.cu:15-61); mask arguments may be None."""
    T = f64(T)
    out = np.empty_like(T)
    lib().or_kappa(ptr(out), ptr(f64(msat0)), ptr(f64(msat0T0)), ptr(T), ptr(f64(Tc)), ptr(f64(S)), ptr(f64(n)),
                   msat0Mul, msat0T0Mul, TcMul, SMul, T.size)
    return out


def ts(N, msat, msat0T0, Tc, S, msatMul, msat0T0Mul, TcMul, SMul):
    """tsAsync (Ts.cu:84-120); mask arguments may be None, N cells."""
    out = np.empty(N)
    lib().or_ts(ptr(out), ptr(f64(msat)), ptr(f64(msat0T0)), ptr(f64(Tc)), ptr(f64(S)), msatMul, msat0T0Mul,
                TcMul, SMul, N)
    return out


def cp(T, Td, n, TdMul):
    """cpAsync (Cp.cu:14-83)."""
    T = f64(T)
    out = np.empty_like(T)
    lib().or_cp(ptr(out), ptr(T), ptr(f64(Td)), ptr(f64(n)), TdMul, T.size)
    return out


def energy_flow(m, R, Tc, S, n, SMul):
    """energyFlowAsync (energy_flow.cu:15-48); m, R of shape (3, N)."""
    m, R = f64(m).reshape(3, -1), f64(R).reshape(3, -1)
    out = np.empty(m.shape[1])
    lib().or_energy_flow(ptr(out), ptr(m[0]), ptr(m[1]), ptr(m[2]), ptr(R[0]), ptr(R[1]), ptr(R[2]),
                         ptr(f64(Tc)), ptr(f64(S)), ptr(f64(n)), SMul, out.size)
    return out


def uniaxial_anisotropy(m, Ku, Ms, K2Mul, u_masks, u_muls):
    """uniaxialAnisotropyAsync (uniaxialanisotropy.cu:12-52)."""
    m = f64(m).reshape(3, -1)
    h = np.empty_like(m)
    um = [f64(x) for x in u_masks]
    lib().or_uniaxial_anisotropy(ptr(h[0]), ptr(h[1]), ptr(h[2]), ptr(m[0]), ptr(m[1]), ptr(m[2]), ptr(f64(Ku)),
                                 ptr(f64(Ms)), K2Mul, ptr(um[0]), u_muls[0], ptr(um[1]), u_muls[1], ptr(um[2]),
                                 u_muls[2], m.shape[1])
    return h


def llbar_nonlocal00nc(H, msat0T0, lam, lmul, N, cell, pbc):
    """llbar_nonlocal00nc_async (llbar-nonlocal00nc.cu:11-208); lam: 3 masks or Nones."""
    Hp, Ha = _planes(H)
    t = np.empty_like(Ha)
    tp = (C.c_void_p * 3)(*[p.ctypes.data for p in t])
    lm = [f64(x) for x in lam]
    lp = (C.c_void_p * 3)(*[ptr(x) for x in lm])
    lmu = (C.c_double * 3)(*lmul)
    lib().or_llbar_nonlocal00nc(tp, Hp, ptr(f64(msat0T0)), lp, lmu, int(N[0]), int(N[1]), int(N[2]),
                                float(cell[0]), float(cell[1]), float(cell[2]), int(pbc[0]), int(pbc[1]),
                                int(pbc[2]))
    return t


def div_mul_pow(a, b, c, p):
    """divMulPowAsync (divMulPow.cu:13-30)."""
    a = f64(a)
    out = np.empty_like(a)
    lib().or_div_mul_pow(ptr(out), ptr(a), ptr(f64(b)), ptr(f64(c)), p, a.size)
    return out


def wavg(a, b, w0, w1, R, w0Mul, w1Mul, RMul):
    """wavgAsync (wavg.cu:13-30)."""
    a = f64(a)
    out = np.empty_like(a)
    lib().or_wavg(ptr(out), ptr(a), ptr(f64(b)), ptr(f64(w0)), ptr(f64(w1)), ptr(f64(R)), w0Mul, w1Mul, RMul,
                  a.size)
    return out
