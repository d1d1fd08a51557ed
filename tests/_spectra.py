"""Spectrum comparison used by the GPU parity tests.

Two numbers per comparison, over the bins whose value is above 1e-12 of the
peak (below that a bin is a handful of terms on the zero crossing of the F
table at x = 20.7, where no float evaluation of x is well conditioned):

 parity   max_e (|gpu - exact| - 2 |exact - fp64|) / |exact|
          exact = the reference's per-eval float terms summed in FP64 (oracle
          EXACT mode), fp64 = the same model in FP64 arithmetic (oracle FP64).
          The reference's own FP32 rounding |exact - fp64| (up to ~1e-5 in
          tail-dominated bins) is not counted against the GPU.
 accuracy max_e |gpu - fp64| / |fp64|

Both must be <= 1e-5.
"""
import numpy as np

TOL = 1e-5


def spectrum_errors(gpu, exact, fp64, floor=1e-12):
    gpu = np.asarray(gpu, np.float64)
    exact = np.asarray(exact, np.float64)
    fp64 = np.asarray(fp64, np.float64)
    peak = np.abs(exact).max()
    live = np.abs(exact) > floor * peak
    if not live.any():
        return 0.0, 0.0, live
    ref_err = np.abs(exact[live] - fp64[live])
    parity = np.max((np.abs(gpu[live] - exact[live]) - 2 * ref_err) / np.abs(exact[live]))
    accuracy = np.max(np.abs(gpu[live] - fp64[live]) / np.abs(fp64[live]))
    return float(max(parity, 0.0)), float(accuracy), live


def oracle_pair(orc, *args, **kw):
    """(exact, fp64) FP64 arrays of the oracle spectrum."""
    _, ex = orc.sync_spectrum(*args, mode=orc.EXACT, want_fp64=True, **kw)
    _, f64 = orc.sync_spectrum(*args, mode=orc.FP64, want_fp64=True, **kw)
    return ex, f64


def assert_spectrum(gpu, exact, fp64, tol=TOL, what=""):
    parity, acc, live = spectrum_errors(gpu, exact, fp64)
    print(f"{what} parity {parity:.2e} accuracy(FP64) {acc:.2e} over {int(live.sum())} bins")
    assert parity <= tol, (what, parity)
    assert acc <= tol, (what, acc)
    peak = np.abs(exact).max()
    assert np.all(np.abs(np.asarray(gpu)[~live]) <= 1e-10 * peak + 1e-37)
