"""Seeded synthetic inputs of the benchmark and test workloads.

All generators use numpy ``default_rng(seed)`` (123 is the seed of the
reference tests, src/tests/synchrotron.py:63) and return float32
structure-of-arrays: ``u`` and ``b`` of shape (3, N).

Momenta: |u| follows a power law dN/d|u| ~ |u|^p on [umin, umax] by inverse
CDF (the form of src/tests/synchrotron.py:65-68), directions isotropic
(cos(theta) ~ U[-1, 1], phi ~ U[0, 2 pi), src/tests/synchrotron.py:72-76).
Field: B = B0 * ((0, 0, 1) + N(0, sigma^2) per component).  E = 0, weights 1.
"""
from __future__ import annotations

import numpy as np


def random_plaw(rng, size, xmin, xmax, p):
    a = xmin ** (p + 1)
    b = xmax ** (p + 1)
    return ((b - a) * rng.random(size) + a) ** (1.0 / (p + 1))


def isotropic(rng, size):
    c = 2.0 * (rng.random(size) - 0.5)
    phi = 2.0 * np.pi * rng.random(size)
    s = np.sqrt(1.0 - c * c)
    return np.stack([s * np.cos(phi), s * np.sin(phi), c])


def particles(n, p=-2.5, umin=1.0, umax=1e3, sigma=0.3, B0=1.0, seed=123):
    """The 2^24-particle emission workload (and its sub/supersets).

    Returns (u, b) float32 arrays of shape (3, n)."""
    rng = np.random.default_rng(seed)
    umag = random_plaw(rng, n, umin, umax, p)
    u = (isotropic(rng, n) * umag).astype(np.float32)
    b = np.zeros((3, n), np.float32)
    b[2] = 1.0
    b += rng.normal(0.0, sigma, size=(3, n)).astype(np.float32)
    if B0 != 1.0:
        b *= np.float32(B0)
    return np.ascontiguousarray(u), np.ascontiguousarray(b)


def cells(ncells=4096, median=4000, sigma=1.0, seed=123, max_total=None):
    """Spatially resolved workload: lognormal counts, per-cell index in
    [-3, -1.5] and field scale log-uniform in [0.1, 10].

    Returns (u, b, offsets int64[ncells + 1], p_cell, b0_cell)."""
    rng = np.random.default_rng(seed)
    counts = np.maximum(1, np.round(rng.lognormal(np.log(median), sigma, ncells))).astype(np.int64)
    if max_total is not None and counts.sum() > max_total:
        counts = np.maximum(1, (counts * (max_total / counts.sum())).astype(np.int64))
    offsets = np.zeros(ncells + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    n = int(offsets[-1])
    p_cell = rng.uniform(-3.0, -1.5, ncells)
    b0_cell = 10.0 ** rng.uniform(-1.0, 1.0, ncells)
    cell_of = np.repeat(np.arange(ncells), counts)
    umag = random_plaw(rng, n, 1.0, 1e3, p_cell[cell_of])
    u = (isotropic(rng, n) * umag).astype(np.float32)
    b = np.zeros((3, n), np.float32)
    b[2] = 1.0
    b += rng.normal(0.0, 0.3, size=(3, n)).astype(np.float32)
    b *= b0_cell[cell_of].astype(np.float32)
    return np.ascontiguousarray(u), np.ascontiguousarray(b), offsets, p_cell, b0_cell


def distributions(nd=65536, ng=512, gmin=1.0, gmax=1e6, seed=123):
    """Batched distribution workload: power law with exponential cutoff,
    index in [-3.5, -1.5], cutoff log-uniform in [1e2, 1e5], each row
    normalised by its trapezoid integral over the log grid given by ``grid``.

    Returns (f float32 (nd, ng), p, gcut)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-3.5, -1.5, nd)
    gcut = 10.0 ** rng.uniform(2.0, 5.0, nd)
    g = np.logspace(np.log10(gmin), np.log10(gmax), ng)
    f = np.empty((nd, ng), np.float32)
    step = 4096
    for s in range(0, nd, step):
        e = slice(s, min(nd, s + step))
        logf = p[e, None] * np.log(g)[None, :] - g[None, :] / gcut[e, None]
        v = np.exp(logf)
        norm = np.trapezoid(v, g, axis=1)
        f[e] = (v / norm[:, None]).astype(np.float32)
    return f, p, gcut


def fit_weights(u, p_theta=-2.2, cut_theta=3e3, p_base=-2.5):
    """Reweighting of a p_base power-law population to index p_theta with an
    exponential cutoff: w = |u|^(p_theta - p_base) exp(-|u| / cut_theta)."""
    umag = np.sqrt((u.astype(np.float64) ** 2).sum(0))
    return (umag ** (p_theta - p_base) * np.exp(-umag / cut_theta)).astype(np.float32)
