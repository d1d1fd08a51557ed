"""ctypes binding of libragnar_hip.so (the C ABI in include/ragnar_hip.h).

Used by the torch autograd path, the tests and the benchmark.  Loading the
library needs no GPU; calling a compute entry point does.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libragnar_hip.so"
HEADER_PATH = _HERE.parent / "include" / "ragnar_hip.h"

RGN_OK, RGN_ERR_INVALID, RGN_ERR_RANGE, RGN_ERR_HIP, RGN_ERR_NOMEM, RGN_ERR_UNSUPPORTED = range(6)
RGN_FTABLE_REFERENCE, RGN_FTABLE_ACCURATE = 0, 1
RGN_PREC_FP32, RGN_PREC_F16X3, RGN_PREC_F16X1, RGN_PREC_BF16X6 = 0, 1, 2, 3
RGN_METHOD_AUTO, RGN_METHOD_DIRECT, RGN_METHOD_BINNED = 0, 1, 2

vp = C.c_void_p
sz = C.c_size_t
fptr3 = C.c_void_p * 3

_SIGNATURES = {
    "rgn_init": (C.c_int, [C.c_int]),
    "rgn_finalize": (C.c_int, []),
    "rgn_is_initialized": (C.c_int, []),
    "rgn_last_error": (C.c_char_p, []),
    "rgn_sync": (C.c_int, [vp]),
    "rgn_version": (C.c_char_p, []),
    "rgn_ftable_build": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(vp)]),
    "rgn_ftable_from_host": (C.c_int, [vp, vp, C.c_int, C.POINTER(vp)]),
    "rgn_ftable_nodes": (C.c_int, [vp, vp, vp]),
    "rgn_ftable_size": (C.c_int, [vp]),
    "rgn_ftable_aux": (C.c_int, [vp, vp, vp]),
    "rgn_ftable_destroy": (C.c_int, [vp]),
    "rgn_ffunc_integrand": (C.c_int, [C.c_float, C.POINTER(C.c_float)]),
    "rgn_bessel_k": (C.c_int, [C.c_double, vp, vp, sz, vp]),
    "rgn_sync_plan_create": (C.c_int, [vp, vp, C.c_int, C.c_float, C.c_float, C.c_float,
                                       C.POINTER(vp)]),
    "rgn_sync_plan_destroy": (C.c_int, [vp]),
    "rgn_sync_spectrum": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, vp, vp]),
    "rgn_sync_adjoint": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, vp, vp, fptr3, fptr3,
                                   fptr3, vp]),
    "rgn_sync_spectrum_method": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, C.c_int, vp, vp]),
    "rgn_sync_adjoint_method": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, vp, C.c_int, vp, fptr3,
                                          fptr3, fptr3, vp]),
    "rgn_sync_cells": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, vp, C.c_int, vp, vp]),
    "rgn_sync_particle_terms": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, vp, vp]),
    "rgn_sync_plan_info": (C.c_int, [vp, vp, vp]),
    "rgn_sync_spectrum_stages": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, C.c_int, vp, vp, vp,
                                           vp, C.c_int, C.POINTER(C.c_int)]),
    "rgn_sync_adjoint_stages": (C.c_int, [vp, fptr3, fptr3, fptr3, vp, sz, vp, C.c_int, vp, fptr3,
                                          fptr3, fptr3, vp, vp, vp, C.c_int, C.POINTER(C.c_int)]),
    "rgn_sync_fromdist": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp]),
    "rgn_ic_spectrum": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, vp]),
    "rgn_sync_fromdist_batched": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, sz, C.c_int, vp, vp]),
    "rgn_sync_fromdist_batched_plaw": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, vp, sz, C.c_int, vp, vp]),
    "rgn_hist_plan_create": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
    "rgn_hist_plan_destroy": (C.c_int, [vp]),
    "rgn_hist_plan_edges": (C.c_int, [vp, vp]),
    "rgn_energy_hist": (C.c_int, [vp, fptr3, sz, C.c_int, vp, vp, vp]),
    "rgn_energy_hist_noalloc": (C.c_int, [vp, fptr3, sz, C.c_int, vp, vp, vp]),
    "rgn_malloc": (C.c_int, [C.POINTER(vp), sz]),
    "rgn_free": (C.c_int, [vp]),
    "rgn_memcpy_h2d": (C.c_int, [vp, vp, sz]),
    "rgn_memcpy_d2h": (C.c_int, [vp, vp, sz]),
    "rgn_memcpy_d2d": (C.c_int, [vp, vp, sz]),
    "rgn_memset": (C.c_int, [vp, C.c_int, sz]),
    "rgn_host_malloc": (C.c_int, [C.POINTER(vp), sz]),
    "rgn_host_free": (C.c_int, [vp]),
    "rgn_memcpy_h2d_async": (C.c_int, [vp, vp, sz, vp]),
    "rgn_stream_create": (C.c_int, [C.POINTER(vp)]),
    "rgn_stream_destroy": (C.c_int, [vp]),
    "rgn_event_create": (C.c_int, [C.POINTER(vp)]),
    "rgn_event_destroy": (C.c_int, [vp]),
    "rgn_event_record": (C.c_int, [vp, vp]),
    "rgn_event_sync": (C.c_int, [vp]),
}


class RagnarError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rgn error {code}] {msg}")
        self.code = code


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Load libragnar_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with ragnar_amd/build.py")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != RGN_OK:
        msg = lib().rgn_last_error().decode(errors="replace")
        if status == RGN_ERR_RANGE:
            raise ValueError(msg)
        raise RagnarError(status, msg)


def declared_symbols() -> list[str]:
    """Every function declared in include/ragnar_hip.h."""
    text = HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgn_[a-z0-9_]+)\s*\(", text)))


def ptr3(*args):
    """Three device addresses (ints or None), given as three arguments or one
    sequence, as a C array of void*."""
    vals = list(args[0]) if len(args) == 1 else list(args)
    if len(vals) != 3:
        raise ValueError("ptr3 needs three addresses")
    return fptr3(*[v if v else None for v in vals])


class FTable:
    """An F(x) table handle."""

    def __init__(self, mode: int = RGN_FTABLE_REFERENCE, npoints: int = 200, xmin: float = 1e-6,
                 xmax: float = 100.0, *, nodes=None):
        h = vp()
        if nodes is None:
            check(lib().rgn_ftable_build(mode, npoints, xmin, xmax, C.byref(h)))
        else:
            import numpy as np

            x = np.ascontiguousarray(nodes[0], np.float32)
            y = np.ascontiguousarray(nodes[1], np.float32)
            check(lib().rgn_ftable_from_host(x.ctypes.data, y.ctypes.data, len(x), C.byref(h)))
        self.handle = h

    def nodes(self):
        import numpy as np

        n = lib().rgn_ftable_size(self.handle)
        x = np.zeros(n, np.float32)
        y = np.zeros(n, np.float32)
        check(lib().rgn_ftable_nodes(self.handle, x.ctypes.data, y.ctypes.data))
        return x, y

    def aux(self):
        import numpy as np

        n = lib().rgn_ftable_size(self.handle)
        k = np.zeros(n)
        d = np.zeros(n)
        check(lib().rgn_ftable_aux(self.handle, k.ctypes.data, d.ctypes.data))
        return k, d

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rgn_ftable_destroy(self.handle)
            self.handle = None


class SyncPlan:
    """Photon-bin plan for the emission kernels."""

    def __init__(self, table: FTable, bins, B0: float = 1.0, g_syn: float = 1.0,
                 e_at: float = 1.0):
        import numpy as np

        self.table = table
        self.bins = np.ascontiguousarray(bins, np.float32)
        h = vp()
        check(lib().rgn_sync_plan_create(table.handle, self.bins.ctypes.data, len(self.bins),
                                         B0, g_syn, e_at, C.byref(h)))
        self.handle = h
        self.nbins = len(self.bins)

    def info(self):
        ints = (C.c_int * 6)()
        dbl = (C.c_double * 3)()
        check(lib().rgn_sync_plan_info(self.handle, ints, dbl))
        return {"Z": ints[0], "TP": ints[1], "NB": ints[2], "M": ints[3], "NK": ints[4],
                "binned": bool(ints[5]), "K": dbl[0], "Amin": dbl[1], "Amax": dbl[2]}

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rgn_sync_plan_destroy(self.handle)
            self.handle = None


class HistPlan:
    """Energy-bin plan for the histogram kernel."""

    def __init__(self, bins, log_bins: bool):
        import numpy as np

        self.bins = np.ascontiguousarray(bins, np.float32)
        h = vp()
        check(lib().rgn_hist_plan_create(self.bins.ctypes.data, len(self.bins), int(log_bins),
                                         C.byref(h)))
        self.handle = h
        self.nb = len(self.bins)

    def edges(self):
        import numpy as np

        e = np.zeros(self.nb + 1)
        check(lib().rgn_hist_plan_edges(self.handle, e.ctypes.data))
        return e

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rgn_hist_plan_destroy(self.handle)
            self.handle = None
