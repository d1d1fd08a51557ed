"""Torch-facing wrappers of the C ABI: device tensors in, device tensors out.

Tensors are PyTorch-ROCm ``cuda`` tensors used only as HBM buffers and
streams; every computation runs in the HIP kernels of libragnar_hip.so on the
current torch stream.  Particle data is structure-of-arrays: ``u``, ``b`` and
``e`` are (3, N) float32 tensors (one contiguous row per component).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi
from ._capi import FTable, HistPlan, SyncPlan, check, lib, ptr3

__all__ = ["FTable", "SyncPlan", "HistPlan", "sync_spectrum", "sync_adjoint", "sync_cells",
           "sync_fromdist", "sync_fromdist_batched", "sync_fromdist_batched_plaw", "energy_hist", "particle_terms",
           "spectrum_stages", "adjoint_stages", "SynchrotronSpectrumFn", "synchrotron_spectrum"]

_default_table: FTable | None = None


def default_table() -> FTable:
    """The reference F table (TabulateFfunc defaults), built once."""
    global _default_table
    if _default_table is None:
        _default_table = FTable(_capi.RGN_FTABLE_REFERENCE, 200, 1e-6, 100.0)
    return _default_table


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _soa(t: torch.Tensor | None, name: str, n: int | None = None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (the kernels run on the GPU only)")
    if t.dtype != torch.float32:
        raise ValueError(f"{name} must be float32")
    if t.dim() != 2 or t.shape[0] != 3:
        raise ValueError(f"{name} must have shape (3, N)")
    if n is not None and t.shape[1] != n:
        raise ValueError(f"{name} has {t.shape[1]} particles, expected {n}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return ptr3([t[k].data_ptr() for k in range(3)])


def _vec(t: torch.Tensor | None, name: str, n: int):
    if t is None:
        return None
    if not t.is_cuda or t.dtype != torch.float32 or t.dim() != 1 or t.shape[0] != n:
        raise ValueError(f"{name} must be a float32 device vector of length {n}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def _mat(t: torch.Tensor, name: str, ncols: int):
    """A row-major (rows, ncols) float32 device matrix."""
    if not t.is_cuda or t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != ncols:
        raise ValueError(f"{name} must be a float32 device tensor of shape (rows, {ncols})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def _out(t: torch.Tensor | None, shape, device: torch.device, name: str = "out") -> torch.Tensor:
    """The caller's output tensor, checked, or a new one: the kernels write
    exactly prod(shape) floats through its pointer."""
    shape = tuple(int(s) for s in shape)
    if t is None:
        return torch.empty(shape, dtype=torch.float32, device=device)
    if (not t.is_cuda or t.dtype != torch.float32 or tuple(t.shape) != shape
            or not t.is_contiguous() or t.device != device):
        raise ValueError(f"{name} must be a contiguous float32 tensor of shape {shape} on {device}")
    return t


METHODS = {"auto": _capi.RGN_METHOD_AUTO, "direct": _capi.RGN_METHOD_DIRECT,
           "binned": _capi.RGN_METHOD_BINNED}


def sync_spectrum(plan: SyncPlan, u, b, w=None, e=None, out=None, method="auto") -> torch.Tensor:
    n = u.shape[1]
    pu = _soa(u, "u")
    out = _out(out, (plan.nbins,), u.device)
    check(lib().rgn_sync_spectrum_method(plan.handle, pu,
                                         _soa(e, "e", n) or ptr3(None, None, None),
                                         _soa(b, "b", n), _vec(w, "w", n), n, METHODS[method],
                                         out.data_ptr(), _stream()))
    return out


def sync_adjoint(plan: SyncPlan, u, b, w, e, gbar, need=("w", "u", "b", "e"), method="auto"):
    n = u.shape[1]
    dev = u.device
    gw = torch.empty(n, dtype=torch.float32, device=dev) if "w" in need else None
    gu = torch.empty(3, n, dtype=torch.float32, device=dev) if "u" in need else None
    gb = torch.empty(3, n, dtype=torch.float32, device=dev) if "b" in need else None
    ge = torch.empty(3, n, dtype=torch.float32, device=dev) if "e" in need else None
    gbar = gbar.contiguous().float()
    none3 = ptr3(None, None, None)
    check(lib().rgn_sync_adjoint_method(
        plan.handle, _soa(u, "u"), _soa(e, "e", n) or none3, _soa(b, "b", n), _vec(w, "w", n), n,
        _vec(gbar, "gbar", plan.nbins), METHODS[method],
        gw.data_ptr() if gw is not None else None,
        _soa(gu, "gu", n) or none3, _soa(gb, "gb", n) or none3, _soa(ge, "ge", n) or none3,
        _stream()))
    return gw, gu, gb, ge


def _stage_buffers():
    return (C.c_float * 16)(), (C.c_char_p * 16)(), C.c_int()


def _stage_list(ms, names, k):
    return [(names[i].decode(), float(ms[i])) for i in range(min(k.value, 16))]


def spectrum_stages(plan: SyncPlan, u, b, w=None, e=None, out=None, method="auto"):
    """One forward call timed per kernel launch with HIP events on the current
    stream: [(kernel name, ms), ...] in launch order."""
    n = u.shape[1]
    pu = _soa(u, "u")
    out = _out(out, (plan.nbins,), u.device)
    ms, names, k = _stage_buffers()
    check(lib().rgn_sync_spectrum_stages(plan.handle, pu,
                                         _soa(e, "e", n) or ptr3(None, None, None),
                                         _soa(b, "b", n), _vec(w, "w", n), n, METHODS[method],
                                         out.data_ptr(), _stream(), C.cast(ms, C.c_void_p),
                                         C.cast(names, C.c_void_p), 16, C.byref(k)))
    return _stage_list(ms, names, k)


def adjoint_stages(plan: SyncPlan, u, b, w, e, gbar, need=("w", "u", "b", "e"), method="auto"):
    """One adjoint call timed per kernel launch: [(kernel name, ms), ...]."""
    n = u.shape[1]
    dev = u.device
    gw = torch.empty(n, dtype=torch.float32, device=dev) if "w" in need else None
    gu = torch.empty(3, n, dtype=torch.float32, device=dev) if "u" in need else None
    gb = torch.empty(3, n, dtype=torch.float32, device=dev) if "b" in need else None
    ge = torch.empty(3, n, dtype=torch.float32, device=dev) if "e" in need else None
    gbar = gbar.contiguous().float()
    none3 = ptr3(None, None, None)
    ms, names, k = _stage_buffers()
    check(lib().rgn_sync_adjoint_stages(
        plan.handle, _soa(u, "u"), _soa(e, "e", n) or none3, _soa(b, "b", n), _vec(w, "w", n), n,
        _vec(gbar, "gbar", plan.nbins), METHODS[method],
        gw.data_ptr() if gw is not None else None,
        _soa(gu, "gu", n) or none3, _soa(gb, "gb", n) or none3, _soa(ge, "ge", n) or none3,
        _stream(), C.cast(ms, C.c_void_p), C.cast(names, C.c_void_p), 16, C.byref(k)))
    return _stage_list(ms, names, k)


def sync_cells(plan: SyncPlan, u, b, offsets, w=None, e=None, out=None) -> torch.Tensor:
    """Spectra of ncells particle slices; offsets: host int64 array of ncells + 1."""
    import numpy as np

    offs = np.ascontiguousarray(offsets, np.int64)
    ncells = len(offs) - 1
    n = u.shape[1]
    if ncells < 1:
        raise ValueError("offsets must hold at least two entries")
    if offs[0] < 0 or offs[-1] > n or np.any(np.diff(offs) < 0):
        raise ValueError("cell offsets must be non-decreasing within [0, N]")
    pu = _soa(u, "u")
    out = _out(out, (ncells, plan.nbins), u.device)
    check(lib().rgn_sync_cells(plan.handle, pu, _soa(e, "e", n) or ptr3(None, None, None),
                               _soa(b, "b", n), _vec(w, "w", n), offs.ctypes.data, ncells,
                               out.data_ptr(), _stream()))
    return out


def sync_fromdist(plan: SyncPlan, gbeta, f, dist_log: bool, out=None) -> torch.Tensor:
    ng = gbeta.shape[0]
    pg, pf = _vec(gbeta, "gbeta", ng), _vec(f, "f", ng)
    out = _out(out, (plan.nbins,), gbeta.device)
    check(lib().rgn_sync_fromdist(plan.handle, pg, pf, ng, int(dist_log), out.data_ptr(), _stream()))
    return out


def ic_spectrum(g, fg, g_log: bool, es, fes, eic, out=None) -> torch.Tensor:
    """ICSpectrum (src/physics/ic.cpp:15-46) on device tensors: E_ic^2 f(E_ic)
    at the energies eic (mc^2) of particles (gamma bins g, values fg; g_log:
    log-spaced bins) scattering soft photons (energies es in mc^2, values fes)."""
    ng, nes, neic = g.shape[0], es.shape[0], eic.shape[0]
    args = (_vec(g, "g", ng), _vec(fg, "fg", ng), ng, int(g_log), _vec(es, "es", nes),
            _vec(fes, "fes", nes), nes, _vec(eic, "eic", neic), neic)
    out = _out(out, (neic,), eic.device)
    check(lib().rgn_ic_spectrum(*args, out.data_ptr(), _stream()))
    return out


def sync_fromdist_batched(plan: SyncPlan, gbeta, f2d, dist_log: bool, precision: int,
                          out=None) -> torch.Tensor:
    ng = gbeta.shape[0]
    pg = _vec(gbeta, "gbeta", ng)
    pf = _mat(f2d, "f", ng)
    nd = f2d.shape[0]
    out = _out(out, (nd, plan.nbins), gbeta.device)
    check(lib().rgn_sync_fromdist_batched(plan.handle, pg, ng, int(dist_log), pf, nd, precision,
                                          out.data_ptr(), _stream()))
    return out


def sync_fromdist_batched_plaw(plan: SyncPlan, gbeta, index, cutoff, dist_log: bool, precision: int,
                               out=None) -> torch.Tensor:
    """Spectra of nd power laws with exponential cutoff, f_d(x) = x^index_d
    exp(-x / cutoff_d) normalised by the trapezoid rule over the gbeta grid
    (2 to 2048 points), generated inside the matrix-core kernels (precision
    RGN_PREC_F16X3, _F16X1 or _BF16X6)."""
    ng = gbeta.shape[0]
    nd = index.shape[0]
    pg = _vec(gbeta, "gbeta", ng)
    pi, pc = _vec(index, "index", nd), _vec(cutoff, "cutoff", nd)
    out = _out(out, (nd, plan.nbins), gbeta.device)
    check(lib().rgn_sync_fromdist_batched_plaw(plan.handle, pg, ng, int(dist_log), pi, pc, nd, precision,
                                               out.data_ptr(), _stream()))
    return out


def energy_hist(plan: HistPlan, u, fourvel: bool = True, counts: bool = True):
    n = u.shape[1]
    hist = torch.empty(plan.nb, dtype=torch.float32, device=u.device)
    cnt = torch.empty(plan.nb, dtype=torch.int64, device=u.device) if counts else None
    check(lib().rgn_energy_hist(plan.handle, _soa(u, "u"), n, int(fourvel), hist.data_ptr(),
                                cnt.data_ptr() if cnt is not None else None, _stream()))
    return hist, cnt


def particle_terms(plan: SyncPlan, u, b, w=None, e=None) -> torch.Tensor:
    n = u.shape[1]
    out = torch.empty(n, 5, dtype=torch.float32, device=u.device)
    check(lib().rgn_sync_particle_terms(plan.handle, _soa(u, "u"), _soa(e, "e", n) or ptr3(None, None, None),
                                        _soa(b, "b", n), _vec(w, "w", n), n, out.data_ptr(),
                                        _stream()))
    return out


class SynchrotronSpectrumFn(torch.autograd.Function):
    """S = SynchrotronSpectrum(u, b, w[, e]) with the hand-written adjoint.

    Forward: the gfx950 forward kernel.  Backward: the adjoint kernel, which
    recomputes gamma and the pitch-angle terms of each particle instead of
    storing them; gradients use the table's segment slopes (the exact adjoint
    of the piecewise-linear interpolant that the forward evaluates).
    """

    @staticmethod
    def forward(ctx, u, b, w, e, plan):
        out = sync_spectrum(plan, u.detach(), b.detach(), None if w is None else w.detach(),
                            None if e is None else e.detach())
        ctx.plan = plan
        ctx.save_for_backward(u, b, w, e)
        return out

    @staticmethod
    def backward(ctx, gout):
        u, b, w, e = ctx.saved_tensors
        need = []
        if ctx.needs_input_grad[0]:
            need.append("u")
        if ctx.needs_input_grad[1]:
            need.append("b")
        if ctx.needs_input_grad[2]:
            need.append("w")
        if ctx.needs_input_grad[3]:
            need.append("e")
        gw, gu, gb, ge = sync_adjoint(ctx.plan, u.detach(), b.detach(),
                                      None if w is None else w.detach(),
                                      None if e is None else e.detach(), gout, need=tuple(need))
        return gu, gb, gw, ge, None


def synchrotron_spectrum(u, b, w=None, e=None, *, plan: SyncPlan):
    return SynchrotronSpectrumFn.apply(u, b, w, e, plan)
