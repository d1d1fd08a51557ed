"""rgn_energy_hist_noalloc (the histogram without device allocation) against
rgn_energy_hist, both called through the C ABI.

Counts must equal bit for bit; the float sums (FP64 sums of the same float
terms, rounded once) to 1e-6 relative, and exactly when the two calls use the
same grid.  Two calls in a row share the fixed array: the second, smaller one
must not see the first one's partials.  A plan with 2000 bins takes the cut
grid (2000 bins x 12 bytes x 2048 blocks exceeds the 8 MiB array).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from ragnar_amd import _capi
from ragnar_amd import workloads
from ragnar_amd._capi import check, lib, ptr3

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _call(entry, plan, u, fourvel):
    n = u.shape[1]
    hist = torch.full((plan.nb,), -1.0, dtype=torch.float32, device=DEV)
    cnt = torch.full((plan.nb,), -1, dtype=torch.int64, device=DEV)
    check(getattr(lib(), entry)(plan.handle, ptr3([u[k].data_ptr() for k in range(3)]), n,
                                int(fourvel), hist.data_ptr(), cnt.data_ptr(),
                                torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return hist.double().cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.parametrize("nb,log_bins", [(200, True), (2000, True), (64, False)])
def test_noalloc_matches_allocating_entry(nb, log_bins):
    bins = np.logspace(-2, 3, nb).astype(np.float32) if log_bins else \
        np.linspace(0.01, 1e3, nb).astype(np.float32)
    plan = _capi.HistPlan(bins, log_bins)
    for n, seed, fourvel in ((3_000_001, 1, True), (20_000, 2, False)):
        u, _ = workloads.particles(n, seed=seed)
        U = torch.from_numpy(u).to(DEV)
        h_ref, c_ref = _call("rgn_energy_hist", plan, U, fourvel)
        h_na, c_na = _call("rgn_energy_hist_noalloc", plan, U, fourvel)
        assert np.array_equal(c_na, c_ref), (nb, n)
        assert c_na.sum() == n
        live = h_ref != 0
        assert np.array_equal(h_na == 0, ~live)
        rel = np.abs(h_na[live] / h_ref[live] - 1).max() if live.any() else 0.0
        print(f"nb={nb} n={n}: counts equal, sums max rel diff {rel:.2e}")
        assert rel <= 1e-6
        if nb * 12 * 2048 <= 8 << 20:
            assert np.array_equal(h_na, h_ref)


def test_noalloc_rejects_bad_arguments():
    plan = _capi.HistPlan(np.logspace(-2, 3, 10).astype(np.float32), True)
    st = lib().rgn_energy_hist_noalloc(None, ptr3([None, None, None]), 10, 1, None, None, None)
    assert st == _capi.RGN_ERR_INVALID
    st = lib().rgn_energy_hist_noalloc(plan.handle, ptr3([None, None, None]), 10, 1, None, None, None)
    assert st == _capi.RGN_ERR_INVALID
    assert b"null" in lib().rgn_last_error()
