"""Inverse-Compton spectrum (ICSpectrum, src/physics/ic.cpp:15-46) on the GPU.

- parity: rgn_ic_spectrum against the oracle's per-triple FP32 terms summed in
  FP64 (the oracle itself is pinned in test_oracle.py by the Thomson and
  Klein-Nishina energy-loss limits and the kinematic cutoff), on the
  reference test's configuration, on linear gamma bins and on broad power laws;
- the reference's own test (src/tests/ic.py:14-50) through the drop-in module;
- the unit checks of the reference kernel's constructor (ic.hpp:52-59);
- bitwise reproducibility (fixed-order FP64 sums).
"""
import numpy as np
import pytest
import torch

from ragnar_amd import torch_ops as T

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(DEV)


def _gpu_ic(g, fg, g_log, es, fes, E):
    return T.ic_spectrum(dev(g), dev(fg), g_log, dev(es), dev(fes), dev(E)).cpu().numpy()


def _cases(orc):
    p, e_break = 1.5, 1e-8
    g = orc.logspace(1e3, 1e7, 200)
    es = orc.logspace(1e-11, 1e-7, 200)
    yield ("reference test", g, orc.generator("plaw", (-p, 1e3, 1e7), g), True, es,
           orc.generator("delta", (e_break, e_break / 10), es), orc.logspace(1e3, 1e7, 200))
    gl = orc.linspace(2.0, 5e3, 300)
    es2 = orc.logspace(1e-5, 1.0, 120)
    yield ("linear gamma bins, Klein-Nishina", gl, orc.generator("plaw", (-2.2, 2.0, 5e3), gl), False, es2,
           orc.generator("plaw", (-1.5, 1e-5, 1.0), es2), orc.logspace(1e-4, 1e4, 257))
    g3 = orc.logspace(1.0, 1e6, 512)
    es3 = orc.logspace(1e-9, 1e-3, 64)
    yield ("broad power laws", g3, orc.generator("bplaw", (1e3, -1.5, -3.2, 1.0, 1e6), g3), True, es3,
           orc.generator("plaw", (-0.5, 1e-9, 1e-3), es3), orc.logspace(1e-6, 1e6, 300))
    # more output bins than the device has workgroup slots per chunk: one chunk
    # of all 1500 gamma bins, staged through LDS in two passes
    g4 = orc.logspace(10.0, 1e6, 1500)
    es4 = orc.logspace(1e-8, 1e-4, 8)
    yield ("one chunk, two LDS stages", g4, orc.generator("plaw", (-2.0, 10.0, 1e6), g4), True, es4,
           orc.generator("plaw", (-1.0, 1e-8, 1e-4), es4), orc.logspace(1e-4, 1e6, 2100))


def test_ic_vs_oracle(orc):
    for name, g, fg, g_log, es, fes, E in _cases(orc):
        assert fg is not None and fes is not None, name
        S = _gpu_ic(g, fg, g_log, es, fes, E)
        _, ex = orc.ic_spectrum(g, fg, g_log, es, fes, E, want_fp64=True)
        _, f64 = orc.ic_spectrum(g, fg, g_log, es, fes, E, mode=orc.FP64, want_fp64=True)
        live = ex > 1e-6 * ex.max()
        err = float(np.max(np.abs(S[live] / ex[live] - 1)))
        err64 = float(np.max(np.abs(S[live] / f64[live] - 1)))
        print(f"{name}: max rel err vs reference terms {err:.2e}, vs FP64 {err64:.2e}")
        assert err < 1e-5, name
        # the kinematic cutoff and the empty bins are exact zeros on both sides
        assert np.array_equal(S[ex == 0], ex[ex == 0]), name


def test_ic_reference_test_through_module():
    """src/tests/ic.py:14-50, run on the drop-in module."""
    from ragnar_amd import ragnar as rg

    def check_slope(xs, ys, p, xmin, xmax, atol=1e-2):
        mask = (xs > xmin) & (xs < xmax)
        pfit, _ = np.polyfit(np.log10(xs[mask]), np.log10(ys[mask]), 1)
        return np.isclose(p, pfit, atol=atol)

    e_break = 1e-8
    p = 1.5
    dist_prtls = rg.TabulatedDistribution(rg.Logbins(1e3, 1e7, 200), rg.PlawGenerator(-p, 1e3, 1e7))
    dist_soft_photons = rg.TabulatedDistribution(rg.Logbins(1e-11, 1e-7, 200, rg.EnergyUnits.mec2),
                                                 rg.DeltaGenerator(e_break, e_break / 10))
    x_prtls = dist_prtls.EnergyBins().as_array()
    y_prtls = dist_prtls.F().as_array()
    assert check_slope(x_prtls, y_prtls, -p, 1e3, 1e7)
    bins_eic = rg.Bins(rg.Logspace(1e3, 1e7, 200), rg.EnergyUnits.mec2)
    eic_2_f_ic = rg.ICSpectrum(dist_prtls, dist_soft_photons, bins_eic)
    x_ic, y_ic = bins_eic.as_array(), eic_2_f_ic.as_array()
    assert check_slope(x_ic, y_ic, -p / 2 + 3 / 2, 2e3, 2e4)
    assert np.isclose(x_ic[np.argmax(y_ic)], e_break * 1e7 ** 2, rtol=0.1)


def test_ic_unit_checks():
    from ragnar_amd import ragnar as rg

    dist = rg.TabulatedDistribution(rg.Logbins(1e3, 1e7, 20), rg.PlawGenerator(-2.0, 1e3, 1e7))
    soft_ok = rg.TabulatedDistribution(rg.Logbins(1e-11, 1e-7, 20, rg.EnergyUnits.mec2),
                                       rg.DeltaGenerator(1e-8, 1e-9))
    soft_ev = rg.TabulatedDistribution(rg.Logbins(1e-11, 1e-7, 20, rg.EnergyUnits.eV),
                                       rg.DeltaGenerator(1e-8, 1e-9))
    with pytest.raises(RuntimeError, match="Soft photons energy bins must be in units of mec"):
        rg.ICSpectrum(dist, soft_ev, rg.Logbins(1e3, 1e7, 20, rg.EnergyUnits.mec2))
    with pytest.raises(RuntimeError, match="E_ic must be in units of mec"):
        rg.ICSpectrum(dist, soft_ok, rg.Logbins(1e3, 1e7, 20))


def test_ic_reproducible_and_empty(orc):
    name, g, fg, g_log, es, fes, E = next(iter(_cases(orc)))
    a = _gpu_ic(g, fg, g_log, es, fes, E)
    b = _gpu_ic(g, fg, g_log, es, fes, E)
    assert np.array_equal(a, b)
    # no particles: zeros; no output bins: nothing to do
    z = T.ic_spectrum(dev(np.zeros(0)), dev(np.zeros(0)), True, dev(es), dev(fes), dev(E)).cpu().numpy()
    assert np.count_nonzero(z) == 0
    e0 = T.ic_spectrum(dev(g), dev(fg), True, dev(es), dev(fes), dev(np.zeros(0)))
    assert e0.numel() == 0
