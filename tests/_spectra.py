"""Spectrum comparison used by the GPU parity tests.

Per photon bin above 1e-12 of the peak (below that a bin is a handful of
terms on the zero crossing of the F table at x = 20.7, where no float
evaluation of x is well conditioned), two relative errors:

 parity    |gpu - exact| / |exact|, exact = the reference's per-eval float
           terms summed in FP64 (oracle EXACT mode);
 accuracy  |gpu - fp64| / |fp64|, fp64 = the same model evaluated in FP64
           (oracle FP64 mode).

Both must be <= 1e-5 in every bin where the reference's own FP32 result
agrees with the FP64 model to 1e-5 (|exact - fp64| / |exact| <= 1e-5).
Where it does not (a tail bin dominated by a few particles whose float peak
energy puts them next to the table's zero crossing), the two references are
further apart than the tolerance, and the GPU must match one of them to
1e-5: the kernels reproduce the reference's float e_peak and chiR bit for bit,
so they may follow either.  Those bins are printed with both errors.
"""
import numpy as np

TOL = 1e-5


def spectrum_errors(gpu, exact, fp64, floor=1e-12, tol=TOL):
    """(parity, accuracy, live mask, indices of the bins where the references
    disagree, the worst min(parity, accuracy) over those bins)."""
    gpu = np.asarray(gpu, np.float64)
    exact = np.asarray(exact, np.float64)
    fp64 = np.asarray(fp64, np.float64)
    peak = np.abs(exact).max()
    live = np.abs(exact) > floor * peak
    if not live.any():
        return 0.0, 0.0, live, np.zeros(0, int), 0.0
    idx = np.flatnonzero(live)
    par = np.abs(gpu[idx] - exact[idx]) / np.abs(exact[idx])
    acc = np.abs(gpu[idx] - fp64[idx]) / np.abs(fp64[idx])
    own = np.abs(exact[idx] - fp64[idx]) / np.abs(exact[idx]) > tol
    parity = float(par[~own].max()) if (~own).any() else 0.0
    accuracy = float(acc[~own].max()) if (~own).any() else 0.0
    either = float(np.minimum(par, acc)[own].max()) if own.any() else 0.0
    return parity, accuracy, live, idx[own], either


def oracle_pair(orc, *args, **kw):
    """(exact, fp64) FP64 arrays of the oracle spectrum."""
    _, ex = orc.sync_spectrum(*args, mode=orc.EXACT, want_fp64=True, **kw)
    _, f64 = orc.sync_spectrum(*args, mode=orc.FP64, want_fp64=True, **kw)
    return ex, f64


def assert_spectrum(gpu, exact, fp64, tol=TOL, what=""):
    parity, acc, live, own, either = spectrum_errors(gpu, exact, fp64, tol=tol)
    print(f"{what} parity {parity:.2e} accuracy(FP64) {acc:.2e} over {int(live.sum())} bins")
    if len(own):
        g = np.asarray(gpu, np.float64)[own]
        ex = np.asarray(exact, np.float64)[own]
        f = np.asarray(fp64, np.float64)[own]
        print(f"  {len(own)} bins where the reference's FP32 result and the FP64 model differ by more "
              f"than {tol:g} (bin: |gpu-exact|/exact, |gpu-fp64|/fp64, |exact-fp64|/exact):",
              ", ".join(f"{i}: {abs(a - b) / abs(b):.1e}, {abs(a - c) / abs(c):.1e}, {abs(b - c) / abs(b):.1e}"
                        for i, a, b, c in zip(own, g, ex, f)))
    assert parity <= tol, (what, parity)
    assert acc <= tol, (what, acc)
    assert either <= tol, (what, either)
    peak = np.abs(exact).max()
    assert np.all(np.abs(np.asarray(gpu)[~live]) <= 1e-10 * peak + 1e-37)
