"""Complex-step reference for the adjoint tests.

The FP64 spectrum model of the oracle (log-linear interpolant between the
float table nodes) written with numpy's complex functions, per particle, so
that Im L_p(theta + i h) / h is the exact derivative of each particle's loss
term without any of the adjoint's formulas.
"""
import numpy as np


def loss_per_particle_complex(u, e, b, w, bins, fx, fy, gbar, B0):
    """L_p = sum_e gbar_e S_pe of each particle, in complex arithmetic: the FP64
    spectrum model (log-linear interpolant between the float table nodes,
    synchrotron.hpp:145-172, tabulation.hpp:19-42) written with numpy's
    complex functions, the segment picked from the real part."""
    g = np.sqrt(1 + (u * u).sum(0))
    beta = u / g
    V = e + np.cross(beta, b, axis=0)
    q = (V * V).sum(0) - ((beta * e).sum(0)) ** 2
    chi = np.sqrt(q) / B0
    eps = g * g * chi
    x = bins.astype(float)[:, None] / eps[None, :]
    lx = np.log(x)
    ln = np.log(fx.astype(float))
    fyd = fy.astype(float)
    i = np.clip(np.searchsorted(ln, lx.real, side="right") - 1, 0, len(fx) - 2)
    F = fyd[i] + (lx - ln[i]) * (fyd[i + 1] - fyd[i]) / (ln[i + 1] - ln[i])
    inside = (x.real >= fx.min()) & (x.real < fx.max())
    F = np.where(inside, F, 0)
    return w * chi * (gbar[:, None] * bins.astype(float)[:, None] * F).sum(0)


def complex_step_grads(u, e, b, w, bins, fx, fy, gbar, B0=1.0, h=1e-30):
    """Per-particle (dL/dw [N], dL/du [3, N], dL/dB [3, N], dL/dE [3, N]) of
    L = sum_e gbar_e S_e by complex steps; u, e, b (3, N) and w (N) are float
    arrays (e may be None)."""
    n = u.shape[1]
    if e is None:
        e = np.zeros((3, n), np.float32)
    args = [np.asarray(a).astype(np.complex128) for a in (u, e, b, w)]
    gbar = np.asarray(gbar, np.float64)

    def grad(idx, c=None):
        a = [x.copy() for x in args]
        if c is None:
            a[idx] += 1j * h
        else:
            a[idx][c] += 1j * h
        return loss_per_particle_complex(a[0], a[1], a[2], a[3], bins, fx, fy, gbar, B0).imag / h

    gw = grad(3)
    gu = np.array([grad(0, c) for c in range(3)])
    ge = np.array([grad(1, c) for c in range(3)])
    gb = np.array([grad(2, c) for c in range(3)])
    return gw, gu, gb, ge
