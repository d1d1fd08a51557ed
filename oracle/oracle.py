"""ORACLE -- test infrastructure only (ctypes front end of oracle/liboracle.so).

The CPU restatement of the reference hot path, used as the checker by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The product
(ragnar_amd/) never imports this module.

Every function here forwards to oracle.cpp, whose header lists the reference
file:line each part follows.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

FAITHFUL, EXACT, FP64 = 0, 1, 2

_f32 = np.ctypeslib.ndpointer(np.float32, flags="C")
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
_u64 = np.ctypeslib.ndpointer(np.uint64, flags="C")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        L.orc_num_threads.restype = C.c_int
        L.orc_logspace.argtypes = [C.c_float, C.c_float, C.c_size_t, _f32]
        L.orc_linspace.argtypes = [C.c_float, C.c_float, C.c_size_t, _f32]
        L.orc_ffunc_integrand.restype = C.c_float
        L.orc_ffunc_integrand.argtypes = [C.c_float]
        L.orc_tabulate_ffunc.argtypes = [C.c_size_t, C.c_float, C.c_float, _f32, _f32]
        L.orc_interp.restype = C.c_float
        L.orc_interp.argtypes = [C.c_float, _f32, _f32, C.c_size_t, C.c_float]
        L.orc_omega_sync.argtypes = [_f32, C.c_float, C.c_float, C.c_float,
                                     C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_omega_sync_fp64.argtypes = [_f64, C.c_double, C.c_double, C.c_double,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_sync_spectrum.argtypes = [C.c_int] + [vp] * 10 + [C.c_size_t, _f32, C.c_size_t,
                                                                _f32, _f32, C.c_size_t,
                                                                C.c_float, C.c_float, C.c_float,
                                                                _f32, vp]
        L.orc_sync_fromdist.argtypes = [C.c_int, _f32, _f32, C.c_size_t, C.c_int, _f32,
                                        C.c_size_t, _f32, _f32, C.c_size_t, C.c_float,
                                        C.c_float, _f32, vp]
        L.orc_sync_adjoint_fp64.argtypes = ([vp] * 10 + [C.c_size_t, _f32, C.c_size_t, _f32, _f32,
                                              C.c_size_t, C.c_double, C.c_double, C.c_double,
                                              _f64, _f64, _f64, _f64, _f64, _f64])
        L.orc_ic_spectrum.argtypes = [C.c_int, _f32, _f32, C.c_size_t, C.c_int, _f32, _f32,
                                      C.c_size_t, _f32, C.c_size_t, _f32, vp]
        L.orc_energy_hist.argtypes = [C.c_int, _f32, _f32, _f32, C.c_size_t, _f32, C.c_size_t,
                                      C.c_int, C.c_int, _f32, vp, vp]
        L.orc_generator.restype = C.c_int
        L.orc_generator.argtypes = [C.c_int, _f32, _f32, C.c_size_t, _f32]
        for name in ("orc_k53_fp64", "orc_F_fp64", "orc_dFdx_fp64"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.orc_bessel_k_libstdcxx.restype = C.c_double
        L.orc_bessel_k_libstdcxx.argtypes = [C.c_double, C.c_double]
        _lib = L
    return _lib


def num_threads() -> int:
    return lib().orc_num_threads()


def logspace(start, stop, num) -> np.ndarray:
    out = np.zeros(num, np.float32)
    lib().orc_logspace(np.float32(start), np.float32(stop), num, out)
    return out


def linspace(start, stop, num) -> np.ndarray:
    out = np.zeros(num, np.float32)
    lib().orc_linspace(np.float32(start), np.float32(stop), num, out)
    return out


def ffunc_integrand(x) -> np.float32:
    return np.float32(lib().orc_ffunc_integrand(np.float32(x)))


def ftable(n=200, xmin=1e-6, xmax=100.0):
    """TabulateFfunc(n, xmin, xmax): (x, y) float32."""
    x = np.zeros(n, np.float32)
    y = np.zeros(n, np.float32)
    lib().orc_tabulate_ffunc(n, np.float32(xmin), np.float32(xmax), x, y)
    return x, y


def interp(x0, x, y, yfill=0.0) -> np.float32:
    """InterpolateTabulatedFunction<true> on log-spaced nodes (tabulation.hpp:19-42)."""
    return np.float32(lib().orc_interp(np.float32(x0), np.ascontiguousarray(x, np.float32),
                                       np.ascontiguousarray(y, np.float32), len(x), np.float32(yfill)))


def omega_sync(u, e, b, B0=1.0, g_syn=1.0, e_at=1.0):
    """(e_peak, chiR) float32 for one particle (reference promotions)."""
    a = np.array([*u, *e, *b], np.float32)
    ep, chi = C.c_float(), C.c_float()
    lib().orc_omega_sync(a, B0, g_syn, e_at, C.byref(ep), C.byref(chi))
    return np.float32(ep.value), np.float32(chi.value)


def omega_sync_fp64(u, e, b, B0=1.0, g_syn=1.0, e_at=1.0):
    a = np.array([*u, *e, *b], np.float64)
    ep, chi = C.c_double(), C.c_double()
    lib().orc_omega_sync_fp64(a, B0, g_syn, e_at, C.byref(ep), C.byref(chi))
    return ep.value, chi.value


def _ptr(a):
    return None if a is None else a.ctypes.data


def _soa(a, n):
    if a is None:
        return [None, None, None]
    a = np.ascontiguousarray(a, np.float32)
    assert a.shape == (3, n)
    return [a[0], a[1], a[2]]


def sync_spectrum(u, b, bins, e=None, w=None, B0=1.0, g_syn=1.0, e_at=1.0, mode=EXACT,
                  table=None, want_fp64=False):
    """SynchrotronSpectrum<3> on SoA (3, N) float32 arrays; returns float32 [nbins]
    (and the FP64 sum for EXACT / FP64 modes when want_fp64)."""
    u = np.ascontiguousarray(u, np.float32)
    n = u.shape[1]
    uu, ee, bb = _soa(u, n), _soa(e, n), _soa(b, n)
    ww = None if w is None else np.ascontiguousarray(w, np.float32)
    bins = np.ascontiguousarray(bins, np.float32)
    fx, fy = table if table is not None else ftable()
    fx = np.ascontiguousarray(fx, np.float32)
    fy = np.ascontiguousarray(fy, np.float32)
    out = np.zeros(len(bins), np.float32)
    outd = np.zeros(len(bins), np.float64)
    lib().orc_sync_spectrum(mode, *[_ptr(c) for c in uu], *[_ptr(c) for c in ee],
                            *[_ptr(c) for c in bb], _ptr(ww), n, bins, len(bins), fx, fy, len(fx),
                            B0, g_syn, e_at, out, outd.ctypes.data)
    return (out, outd) if want_fp64 else out


def sync_adjoint_fp64(u, b, bins, gbar, e=None, w=None, B0=1.0, g_syn=1.0, e_at=1.0, table=None):
    """FP64 gradient of L = sum_e gbar[e] S[e] (the ORC_FP64 model) per particle:
    (gw [N], gu [3, N], gb [3, N], ge [3, N], dnode [N]); dnode is the smallest
    distance of any of the particle's x_e from a table node, in segments."""
    u = np.ascontiguousarray(u, np.float32)
    n = u.shape[1]
    uu, ee, bb = _soa(u, n), _soa(e, n), _soa(b, n)
    ww = None if w is None else np.ascontiguousarray(w, np.float32)
    bins = np.ascontiguousarray(bins, np.float32)
    fx, fy = table if table is not None else ftable()
    gw, dn = np.zeros(n), np.zeros(n)
    gu, gb, ge = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
    lib().orc_sync_adjoint_fp64(*[_ptr(c) for c in uu], *[_ptr(c) for c in ee], *[_ptr(c) for c in bb],
                                _ptr(ww), n, bins, len(bins), np.ascontiguousarray(fx, np.float32),
                                np.ascontiguousarray(fy, np.float32), len(fx), B0, g_syn, e_at,
                                np.ascontiguousarray(gbar, np.float64), gw, gu, gb, ge, dn)
    return gw, gu, gb, ge, dn


def sync_fromdist(gbeta, f, dist_log, bins, g_syn=1.0, e_at=1.0, mode=EXACT, table=None,
                  want_fp64=False):
    gbeta = np.ascontiguousarray(gbeta, np.float32)
    f = np.ascontiguousarray(f, np.float32)
    bins = np.ascontiguousarray(bins, np.float32)
    fx, fy = table if table is not None else ftable()
    out = np.zeros(len(bins), np.float32)
    outd = np.zeros(len(bins), np.float64)
    lib().orc_sync_fromdist(mode, gbeta, f, len(gbeta), int(dist_log), bins, len(bins),
                            np.ascontiguousarray(fx, np.float32), np.ascontiguousarray(fy, np.float32),
                            len(fx), g_syn, e_at, out, outd.ctypes.data)
    return (out, outd) if want_fp64 else out


def ic_spectrum(g, fg, g_log, es, fes, eic, mode=EXACT, want_fp64=False):
    """ICSpectrum (ic.cpp:15-46): E_ic^2 f(E_ic) at eic; (float32, FP64 sums)."""
    g = np.ascontiguousarray(g, np.float32)
    fg = np.ascontiguousarray(fg, np.float32)
    es = np.ascontiguousarray(es, np.float32)
    fes = np.ascontiguousarray(fes, np.float32)
    eic = np.ascontiguousarray(eic, np.float32)
    out = np.zeros(len(eic), np.float32)
    outd = np.zeros(len(eic), np.float64)
    lib().orc_ic_spectrum(mode, g, fg, len(g), int(g_log), es, fes, len(es), eic, len(eic), out,
                          outd.ctypes.data)
    return (out, outd) if want_fp64 else out


def energy_hist(u, bins, log_bins, fourvel=True, mode=EXACT):
    """(hist float32, FP64 sums, uint64 counts)."""
    u = np.ascontiguousarray(u, np.float32)
    bins = np.ascontiguousarray(bins, np.float32)
    n = u.shape[1]
    h = np.zeros(len(bins), np.float32)
    hd = np.zeros(len(bins), np.float64)
    c = np.zeros(len(bins), np.uint64)
    lib().orc_energy_hist(mode, u[0], u[1], u[2], n, bins, len(bins), int(log_bins), int(fourvel),
                          h, hd.ctypes.data, c.ctypes.data)
    return h, hd, c


def generator(kind, params, e):
    """kind: 'plaw' | 'bplaw' | 'delta'; returns f or None when the reference would throw."""
    k = {"plaw": 0, "bplaw": 1, "delta": 2}[kind]
    e = np.ascontiguousarray(e, np.float32)
    f = np.zeros(len(e), np.float32)
    st = lib().orc_generator(k, np.ascontiguousarray(params, np.float32), e, len(e), f)
    return None if st else f


def k53(x) -> float:
    return lib().orc_k53_fp64(float(x))


def bessel_k_libstdcxx(nu, x) -> float:
    return lib().orc_bessel_k_libstdcxx(float(nu), float(x))


def F_fp64(x) -> float:
    return lib().orc_F_fp64(float(x))


def dFdx_fp64(x) -> float:
    return lib().orc_dFdx_fp64(float(x))
