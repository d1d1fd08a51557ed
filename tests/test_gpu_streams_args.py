"""Argument checks of the torch wrappers and the ordering of calls that share
the library's scratch buffers across streams.

Argument errors are raised as ValueError on the host, before any launch.
Calls issued on two side streams without synchronisation between them give
the same results as the same calls issued on one stream.
"""
import numpy as np
import pytest
import torch

from ragnar_amd import torch_ops as T
from ragnar_amd import workloads

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def plan():
    return T.SyncPlan(T.default_table(), np.logspace(-2, 7, 256).astype(np.float32), 1.0, 1.0, 1.0)


def test_output_and_matrix_checks(plan):
    u, b = (torch.from_numpy(a).to(DEV) for a in workloads.particles(1000, seed=3))
    bad_outs = [torch.empty(256),                                   # host tensor
                torch.empty(255, device=DEV),                       # short
                torch.empty(256, dtype=torch.float64, device=DEV),  # dtype
                torch.empty(512, device=DEV)[::2]]                  # strided
    for out in bad_outs:
        with pytest.raises(ValueError):
            T.sync_spectrum(plan, u, b, out=out)
        with pytest.raises(ValueError):
            T.spectrum_stages(plan, u, b, out=out)
    g = torch.logspace(0, 2, 64, device=DEV)
    f = torch.ones(64, device=DEV)
    with pytest.raises(ValueError):
        T.sync_fromdist(plan, g, f, True, out=torch.empty(100, device=DEV))
    with pytest.raises(ValueError):  # host distribution matrix
        T.sync_fromdist_batched(plan, g, torch.ones(8, 64), True, 1)
    with pytest.raises(ValueError):  # wrong column count
        T.sync_fromdist_batched(plan, g, torch.ones(8, 63, device=DEV), True, 1)
    with pytest.raises(ValueError):  # one-dimensional
        T.sync_fromdist_batched(plan, g, torch.ones(64, device=DEV), True, 1)
    with pytest.raises(ValueError):  # output of the wrong shape
        T.sync_fromdist_batched(plan, g, torch.ones(8, 64, device=DEV), True, 1,
                                out=torch.empty(8, 255, device=DEV))
    offs = np.array([0, 400, 1000])
    with pytest.raises(ValueError):
        T.sync_cells(plan, u, b, offs, out=torch.empty(2, 200, device=DEV))
    with pytest.raises(ValueError):
        T.sync_cells(plan, u, b, np.array([0, 600, 500]))
    with pytest.raises(ValueError):
        T.sync_cells(plan, u, b, np.array([0, 1001]))
    e = torch.logspace(-2, 2, 50, device=DEV)
    with pytest.raises(ValueError):
        T.ic_spectrum(g, f, True, e, e, e, out=torch.empty(50))
    # a valid call still runs
    S = T.sync_spectrum(plan, u, b, out=torch.empty(256, device=DEV))
    assert torch.isfinite(S).all()


def test_binned_calls_on_two_streams(plan):
    """Forward and adjoint calls of the factorised method (which share the run
    table, the sorted sources and the moment buffers) alternate between two
    side streams with no synchronisation between them."""
    sets = []
    for seed, n in ((21, 3_000_000), (22, 2_000_000)):
        u, b = (torch.from_numpy(a).to(DEV) for a in workloads.particles(n, seed=seed))
        sets.append((u, b))
    gbar = torch.from_numpy(np.random.default_rng(4).normal(0, 1, 256).astype(np.float32)).to(DEV)
    ref_f, ref_g = [], []
    for u, b in sets:
        ref_f.append(T.sync_spectrum(plan, u, b, method="binned").clone())
        ref_g.append(T.sync_adjoint(plan, u, b, None, None, gbar, need=("w",), method="binned")[0].clone())
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    outs_f, outs_g = [], []
    for rep in range(3):
        for k, st in enumerate((s1, s2)):
            u, b = sets[k]
            with torch.cuda.stream(st):
                outs_f.append((k, T.sync_spectrum(plan, u, b, method="binned")))
                outs_g.append((k, T.sync_adjoint(plan, u, b, None, None, gbar, need=("w",),
                                                 method="binned")[0]))
    torch.cuda.synchronize()
    for k, S in outs_f:
        d = (S - ref_f[k]).abs().max().item() / ref_f[k].abs().max().item()
        assert d <= 1e-6, (k, d)
    for k, gw in outs_g:
        assert torch.equal(gw, ref_g[k]), k
