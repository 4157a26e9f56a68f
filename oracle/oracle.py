"""Test infrastructure only (the oracle): ctypes bindings to

* libfsoracle.so  - the FP64 C restatement of the reference path (fsoracle.c)
* _ref/libfsref.so - the reference's own C++ compiled from its sources by
  oracle/Makefile (present only where it was built)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libfsoracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfsref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_cd, _ci, _cll, _cf = C.c_double, C.c_int, C.c_longlong, C.c_float


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class _Lib:
    prefix = ""

    def __init__(self, path):
        self.path = path
        self.lib = C.CDLL(path)
        p = self.prefix
        L = self.lib
        getattr(L, p + "interpolate").argtypes = [_ci, _ci, _ci, _cd, _cd, _cd, _cd, _cd, _cd, _cd, _cd,
                                                  _f32p, _f32p, _f32p, _f32p, _f32p, _cll, _i32p, _f64p, _ci, _f64p]
        getattr(L, p + "interpolate").restype = _ci
        getattr(L, p + "near_particles").argtypes = [_cd, _f32p, _f32p, _cll, _i32p, _f64p, _ci,
                                                     _i64p, _i32p, _f64p, _cll]
        getattr(L, p + "near_particles").restype = _cll
        getattr(L, p + "near_lines").argtypes = [_cd, _f32p, _f32p, _cll, _i32p, _f64p, _ci, _i32p, _cll]
        getattr(L, p + "near_lines").restype = _cll
        getattr(L, p + "get_near_lines").argtypes = [_cd, _i32p, _f64p, _ci, _f32p, _cf, _i32p, _f64p, _ci]
        getattr(L, p + "get_near_lines").restype = _ci
        getattr(L, p + "w").argtypes = [_cd, _cd, C.POINTER(_cd), C.POINTER(_cd)]
        getattr(L, p + "w").restype = None
        getattr(L, p + "profile").argtypes = [_cd, _cd]
        getattr(L, p + "profile").restype = _cd
        getattr(L, p + "sph_cubic_kernel").argtypes = [_cd]
        getattr(L, p + "sph_cubic_kernel").restype = _cd
        getattr(L, p + "sph_quintic_kernel").argtypes = [_cd]
        getattr(L, p + "sph_quintic_kernel").restype = _cd
        getattr(L, p + "kern_frac").argtypes = [_ci, _cd, _cd, _cd, _cd, _cd]
        getattr(L, p + "kern_frac").restype = _cd
        getattr(L, p + "tau_kern_outer").argtypes = [_cd, _cd, _cd, _cd, _ci, _cd, _cd]
        getattr(L, p + "tau_kern_outer").restype = _cd
        getattr(L, p + "add_tau_particle").argtypes = [_cd, _cd, _cd, _cd, _cd, _cd, _cd, _ci, _cd, _f64p,
                                                       _ci, _cd, _cf, _cf, _cf, _cf, _cf]
        getattr(L, p + "add_tau_particle").restype = None
        getattr(L, p + "add_colden_particle").argtypes = [_cd, _cd, _cd, _cd, _cd, _cd, _cd, _ci, _cd, _f64p,
                                                          _ci, _cd, _cf, _cf, _cf]
        getattr(L, p + "add_colden_particle").restype = None
        getattr(L, p + "set_threads").argtypes = [_ci]
        getattr(L, p + "set_threads").restype = None
        getattr(L, p + "max_threads").argtypes = []
        getattr(L, p + "max_threads").restype = _ci

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    # -- whole path ---------------------------------------------------------
    def interpolate(self, compute_tau, nbins, kernel, box, velfac, atime, lambda_cm, gamma, fosc,
                    amumass, tautail, pos, vel, dens, temp, h, axis, cofm):
        pos = _c(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        axis = _c(axis, np.int32)
        cofm = _c(cofm, np.float64).reshape(-1, 3)
        out = np.zeros((axis.shape[0], nbins), dtype=np.float64)
        vel = _c(vel, np.float32)
        temp = _c(temp, np.float32)
        if compute_tau:
            assert vel.size == 3 * n and temp.size == n
        rc = self._f("interpolate")(int(compute_tau), int(nbins), int(kernel), box, velfac, atime,
                                    lambda_cm, gamma, fosc, amumass, tautail, pos.ravel(), vel.ravel(),
                                    _c(dens, np.float32), temp, _c(h, np.float32), n, axis,
                                    cofm.ravel(), axis.shape[0], out)
        assert rc == 0
        return out

    def near_particles(self, box, pos, h, axis, cofm):
        """Returns (offsets[L+1], parts[M], dr2[M]) in ascending particle order."""
        pos = _c(pos, np.float32).reshape(-1, 3)
        axis = _c(axis, np.int32)
        cofm = _c(cofm, np.float64).reshape(-1, 3)
        L = axis.shape[0]
        off = np.zeros(L + 1, dtype=np.int64)
        dummy_i = np.zeros(1, dtype=np.int32)
        dummy_d = np.zeros(1, dtype=np.float64)
        hh = _c(h, np.float32)
        m = self._f("near_particles")(box, pos.ravel(), hh, pos.shape[0], axis, cofm.ravel(), L,
                                      off, dummy_i, dummy_d, 0)
        parts = np.zeros(max(m, 1), dtype=np.int32)
        dr2 = np.zeros(max(m, 1), dtype=np.float64)
        m2 = self._f("near_particles")(box, pos.ravel(), hh, pos.shape[0], axis, cofm.ravel(), L,
                                       off, parts, dr2, m)
        assert m2 == m
        return off, parts[:m], dr2[:m]

    def table_seconds(self, box, pos, h, axis, cofm):
        """Wall time of building the neighbour lists alone (IndexTable +
        get_near_particles, counted but not copied out), one call."""
        import time
        pos = _c(pos, np.float32).reshape(-1, 3)
        axis = _c(axis, np.int32)
        cofm = _c(cofm, np.float64).reshape(-1, 3)
        L = axis.shape[0]
        off = np.zeros(L + 1, dtype=np.int64)
        dummy_i = np.zeros(1, dtype=np.int32)
        dummy_d = np.zeros(1, dtype=np.float64)
        hh = _c(h, np.float32)
        t0 = time.perf_counter()
        m = self._f("near_particles")(box, pos.ravel(), hh, pos.shape[0], axis, cofm.ravel(), L,
                                      off, dummy_i, dummy_d, 0)
        return time.perf_counter() - t0, int(m)

    def near_lines(self, box, pos, h, axis, cofm):
        pos = _c(pos, np.float32).reshape(-1, 3)
        axis = _c(axis, np.int32)
        cofm = _c(cofm, np.float64).reshape(-1, 3)
        hh = _c(h, np.float32)
        dummy = np.zeros(1, dtype=np.int32)
        n = self._f("near_lines")(box, pos.ravel(), hh, pos.shape[0], axis, cofm.ravel(), axis.shape[0], dummy, 0)
        out = np.zeros(max(n, 1), dtype=np.int32)
        n2 = self._f("near_lines")(box, pos.ravel(), hh, pos.shape[0], axis, cofm.ravel(), axis.shape[0], out, n)
        assert n2 == n
        return out[:n]

    def get_near_lines(self, box, axis, cofm, pos3, hh):
        axis = _c(axis, np.int32)
        cofm = _c(cofm, np.float64).reshape(-1, 3)
        lines = np.zeros(axis.shape[0] + 1, dtype=np.int32)
        dr2 = np.zeros(axis.shape[0] + 1, dtype=np.float64)
        n = self._f("get_near_lines")(box, axis, cofm.ravel(), axis.shape[0], _c(pos3, np.float32), hh,
                                      lines, dr2, lines.shape[0])
        return dict(zip(lines[:n].tolist(), dr2[:n].tolist()))

    # -- pieces ---------------------------------------------------------------
    def w(self, re, im):
        a, b = _cd(), _cd()
        self._f("w")(re, im, C.byref(a), C.byref(b))
        return complex(a.value, b.value)

    def profile(self, u, a):
        return self._f("profile")(u, a)

    def sph_cubic_kernel(self, q):
        return self._f("sph_cubic_kernel")(q)

    def sph_quintic_kernel(self, q):
        return self._f("sph_quintic_kernel")(q)

    def kern_frac(self, kernel, zlow, zhigh, smooth, dr2, zrange):
        return self._f("kern_frac")(kernel, zlow, zhigh, smooth, dr2, zrange)

    def tau_kern_outer(self, btherm, vdr2, vsmooth, aa, kernel, vlow, vhigh):
        return self._f("tau_kern_outer")(btherm, vdr2, vsmooth, aa, kernel, vlow, vhigh)

    def add_tau_particle(self, lam, gamma, fosc, amu, velfac, box, atime, kernel, tautail, tau,
                         dr2, dens, ppos, pvel, temp, smooth):
        self._f("add_tau_particle")(lam, gamma, fosc, amu, velfac, box, atime, kernel, tautail, tau,
                                    tau.shape[0], dr2, dens, ppos, pvel, temp, smooth)

    def add_colden_particle(self, lam, gamma, fosc, amu, velfac, box, atime, kernel, tautail, colden,
                            dr2, dens, ppos, smooth):
        self._f("add_colden_particle")(lam, gamma, fosc, amu, velfac, box, atime, kernel, tautail,
                                       colden, colden.shape[0], dr2, dens, ppos, smooth)

    def set_threads(self, n):
        self._f("set_threads")(int(n))

    def max_threads(self):
        return self._f("max_threads")()


class Ref(_Lib):
    """The reference C++ (compiled from /root/reference by `make -C oracle ref`)."""
    prefix = "fsref_"


class Oracle(_Lib):
    """The FP64 C restatement (oracle/fsoracle.c)."""
    prefix = "fso_"

    def __init__(self, path):
        super().__init__(path)
        L = self.lib
        L.fso_counters.argtypes = [_i64p]
        L.fso_counters.restype = None

    def counters(self):
        """(pairs, window pixels, Voigt samples, colden pixels) of the last call."""
        c = np.zeros(4, dtype=np.int64)
        self.lib.fso_counters(c)
        return {"pairs": int(c[0]), "tau_pixels": int(c[1]), "voigt_samples": int(c[2]),
                "colden_pixels": int(c[3])}


_cache = {}


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def load_ref() -> Ref:
    if "ref" not in _cache:
        _cache["ref"] = Ref(REF_SO)
    return _cache["ref"]


def load_oracle() -> Oracle:
    if "oracle" not in _cache:
        _cache["oracle"] = Oracle(ORACLE_SO)
    return _cache["oracle"]
