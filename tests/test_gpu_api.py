"""The reference's Python test suite (src/tests/*.py), run against the drop-in
module ``ragnar_amd.ragnar`` (same calls and assertions, GPU backend).
"""
import numpy as np
import pytest

from _spectra import assert_spectrum, oracle_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rg():
    from ragnar_amd import ragnar as rg

    rg.Initialize()
    return rg


def test_arrays_bins(rg):
    # src/tests/arrays_bins.py:7-28
    array = np.arange(123)
    assert all(np.isclose(rg.Array1D_i(array.astype(int)).as_array(), array))
    assert all(np.isclose(rg.Array1D_f(array.astype(np.float32)).as_array(), array))
    assert all(np.isclose(rg.Array1D_d(array.astype(np.float64)).as_array(), array))
    assert rg.Array1D_d(array.astype(np.float64)).extent() == 123
    bins = rg.Bins(np.logspace(-2.5, 2.5, 213))
    bins.log_spaced = True
    bins.unit = rg.EnergyUnits.mec2
    assert bins.unit == rg.EnergyUnits.mec2
    assert bins.extent() == 213
    assert bins.log_spaced
    assert repr(bins) == "1D Array [ size: 213 ]"
    with pytest.raises(ValueError):
        bins.head(300)


def test_linspace_logspace(rg):
    # src/tests/linspace_logspace.py:7-36
    assert all(np.isclose(rg.Linspace(-32.0, 4832.0, 56).as_array(), np.linspace(-32.0, 4832.0, 56)))
    assert all(np.isclose(rg.Logspace(10**-2.5, 10**2.5, 213).as_array(), np.logspace(-2.5, 2.5, 213)))
    linbins = rg.Linbins(55.0, 56.0, 123)
    assert all(np.isclose(linbins.as_array(), np.linspace(55.0, 56.0, 123)))
    assert not linbins.log_spaced
    logbins = rg.Logbins(1, 1e3, 22)
    assert all(np.isclose(logbins.as_array(), np.logspace(0, 3, 22)))
    assert logbins.log_spaced
    with pytest.raises(RuntimeError):
        rg.Logspace(0.0, 1.0, 5)
    with pytest.raises(RuntimeError):
        rg.Linspace(2.0, 1.0, 5)


def test_grids_bit_exact_with_oracle(rg, orc):
    for a, b, n in ((1e-6, 100, 200), (1e-2, 1e7, 256), (0.01, 1e5, 200), (1, 1e3, 100)):
        assert np.array_equal(rg.Logspace(a, b, n).as_array(), orc.logspace(a, b, n))
    assert np.array_equal(rg.Linspace(-3.0, 7.5, 31).as_array(), orc.linspace(-3.0, 7.5, 31))


@pytest.mark.parametrize("log", [True, False])
def test_tabulated_function(rg, log):
    """TabulatedFunction / TabulatedFunction_log (tabulation.cpp:19-171): built
    from numpy or from Array1D_f, the six accessors, the MinMax of unsorted x
    (tabulation.cpp:84-102), and the range_error -> ValueError checks of
    tabulation.cpp:104-117 with their messages.  From numpy the reference reads
    x.at(i), y.at(i) for i < len(x) (tabulation.cpp:43-61): a shorter y is
    pybind11's IndexError, a longer y is cut to len(x)."""
    cls = rg.TabulatedFunction_log if log else rg.TabulatedFunction
    x = np.array([3.0, 0.5, 10.0, 2.0, 7.25], np.float32)
    y = np.array([1.0, -2.0, 0.25, 4.0, 8.0], np.float32)
    for tf in (cls(x, y), cls(rg.Array1D_f(x), rg.Array1D_f(y)), cls(x=x, y=y, yfill=0.0)):
        assert np.array_equal(tf.xArr().as_array(), x)
        assert np.array_equal(tf.yArr().as_array(), y)
        assert tf.nPoints() == 5
        assert tf.yFill() == 0.0
        assert tf.xMin() == np.float32(0.5) and tf.xMax() == np.float32(10.0)
    assert cls(x, y, -1.5).yFill() == np.float32(-1.5)
    assert cls(rg.Array1D_f(x), rg.Array1D_f(y), yfill=3.0).yFill() == 3.0
    # a longer y from numpy is cut to len(x)
    tf = cls(x, np.arange(9, dtype=np.float32))
    assert tf.nPoints() == 5 and np.array_equal(tf.yArr().as_array(), np.arange(5, dtype=np.float32))
    with pytest.raises(IndexError, match="index 3 is out of bounds for axis 0 with size 3"):
        cls(x, y[:3])
    # y.size != x.size (Array1D inputs)
    with pytest.raises(ValueError, match=r"^y.size != x.size in TabulatedFunction$"):
        cls(rg.Array1D_f(x), rg.Array1D_f(y[:4]))
    # xmin >= xmax: constant x, one point, and no points
    for xs in (np.full(4, 2.0, np.float32), np.ones(1, np.float32), np.zeros(0, np.float32)):
        with pytest.raises(ValueError, match=r"^xmin >= xmax in TabulatedFunction$"):
            cls(xs, np.zeros(len(xs), np.float32))
    # xmin <= 0 only for the log variant
    xz = np.array([0.0, 1.0, 2.0], np.float32)
    xn = np.array([-1.0, 1.0, 2.0], np.float32)
    yz = np.ones(3, np.float32)
    if log:
        for xs in (xz, xn):
            with pytest.raises(ValueError, match=r"^xmin <= 0.0 in Logspace TabulatedFunction$"):
                cls(xs, yz)
    else:
        assert cls(xn, yz).xMin() == -1.0 and cls(xz, yz).xMin() == 0.0


def _random_components(rng, names):
    return {k: rng.random(10) for k in names}


def test_particles_roundtrip_and_append(rg):
    # src/tests/particles.py:7-96
    prtls = rg.Particles_3D("electrons")
    assert prtls.label() == "electrons"
    assert not prtls.is_allocated()
    rng = np.random.default_rng(123)
    names = ["U1", "U3", "E1", "E2", "E3", "B1", "B2", "B3"]
    first = _random_components(rng, names)
    prtls.fromArrays(first)
    getters = {"U1": lambda: prtls.U(1), "U3": lambda: prtls.U(3), "E1": lambda: prtls.E(1),
               "E2": lambda: prtls.E(2), "E3": lambda: prtls.E(3), "B1": lambda: prtls.B(1),
               "B2": lambda: prtls.B(2), "B3": lambda: prtls.B(3)}
    for k in names:
        assert np.all(np.isclose(first[k], getters[k]().as_array()))
    assert prtls.is_allocated() and prtls.nactive() == 10 and prtls.nalloc() == 10
    second = _random_components(rng, names)
    prtls.fromArrays(second, append=True)
    for k in names:
        got = getters[k]().as_array()
        assert np.all(np.isclose(first[k], got[:10]))
        assert np.all(np.isclose(second[k], got[10:]))
    assert prtls.nactive() == 20 and prtls.nalloc() == 20
    assert len(prtls) == 20
    with pytest.raises(RuntimeError):
        prtls.fromArrays(second)  # already allocated, append not given
    with pytest.raises(IndexError):
        prtls.U(4)


def test_particles_preallocated(rg):
    # src/tests/particles.py:98-171
    rng = np.random.default_rng(123)
    prtls = rg.Particles_3D("electrons")
    prtls.allocate(20)
    assert prtls.nactive() == 0 and prtls.nalloc() == 20
    names = ["U1", "U3", "E1", "E2", "E3", "B1", "B2", "B3"]
    first = _random_components(rng, names)
    prtls.fromArrays(first, append=True)
    assert prtls.nactive() == 10 and prtls.nalloc() == 20
    second = _random_components(rng, names)
    prtls.fromArrays(second, append=True)
    assert prtls.nactive() == 20 and prtls.nalloc() == 20
    for k in names:
        q = {"U": prtls.U, "E": prtls.E, "B": prtls.B}[k[0]](int(k[1])).as_array()
        assert np.all(np.isclose(first[k], q[:10])) and np.all(np.isclose(second[k], q[10:]))


def test_particles_coords_and_missing(rg):
    # src/tests/particles.py:173-211
    rng = np.random.default_rng(123)
    prtls = rg.Particles_3D("electrons")
    X1s = rng.random(10)
    U3s = rng.random(10)
    prtls.fromArrays({"X1": X1s, "U3": U3s}, append=True)
    assert np.all(np.isclose(prtls.X(1).as_array(), X1s))
    assert np.all(np.isclose(prtls.U(3).as_array(), U3s))
    assert np.all(np.isclose(prtls.U(1).as_array(), 0))
    assert np.all(np.isclose(prtls.E(1).as_array(), 0))
    with pytest.raises(RuntimeError):
        rg.Particles_3D("x").fromArrays({"U1": np.ones(3), "U2": np.ones(4)})
    assert "Particles<3D> (electrons)" in repr(prtls)


def test_distributions_normalisation(rg):
    # src/tests/distributions.py:7-29
    for norm in [rg.Logspace, rg.Linspace]:
        bins = rg.Bins(norm(1e-2, 1, int(1e6)))
        dist = rg.TabulatedDistribution(bins, rg.DeltaGenerator(2e-2, 0.01))
        assert np.isclose(np.trapezoid(dist.F().as_array(), dist.EnergyBins().as_array()), 1)
    for norm in [rg.Logspace, rg.Linspace]:
        bins = rg.Bins(norm(1e-3, 10, int(1e6)))
        dist = rg.TabulatedDistribution(bins, rg.BrokenPlawGenerator(0.3, 0.23, -1.0, 1e-2, 2))
        assert np.isclose(np.trapezoid(dist.F().as_array(), dist.EnergyBins().as_array()), 1)
    for norm in [rg.Logspace, rg.Linspace]:
        bins = rg.Bins(norm(1e-2, 1, int(1e6)))
        dist = rg.TabulatedDistribution(bins, rg.PlawGenerator(-1.2, 1e-2, 1))
        assert np.isclose(np.trapezoid(dist.F().as_array(), dist.EnergyBins().as_array()), 1)
    with pytest.raises(RuntimeError):
        rg.PlawGenerator(-2.0)


def test_generators_bit_exact_with_oracle(rg, orc):
    e = orc.logspace(1e-3, 10, 500)
    bins = rg.Bins(e)
    assert np.array_equal(rg.PlawGenerator(-1.7, 1e-2, 5).compute(bins).as_array(),
                          orc.generator("plaw", [-1.7, 1e-2, 5], e))
    assert np.array_equal(rg.BrokenPlawGenerator(0.3, 0.23, -2.5, 1e-2, 2).compute(bins).as_array(),
                          orc.generator("bplaw", [0.3, 0.23, -2.5, 1e-2, 2], e))
    assert np.array_equal(rg.DeltaGenerator(0.5, 0.1).compute(bins).as_array(),
                          orc.generator("delta", [0.5, 0.1], e))


def _check_slope(xs, ys, p, xmin, xmax, atol=1e-2):
    mask = (xs > xmin) & (xs < xmax)
    pfit, _ = np.polyfit(np.log10(xs[mask]), np.log10(ys[mask]), 1)
    return np.isclose(p, pfit, atol=atol)


def test_sync_log(rg):
    # src/tests/synchrotron.py:14-35
    p = 2.23
    prtl_dist = rg.TabulatedDistribution(rg.Logbins(1, 1000, 200), rg.PlawGenerator(-p, 1, 1000))
    esync_bins = rg.Logbins(0.01, 1e7, 200)
    esync_bins.unit = rg.EnergyUnits.mec2
    s = rg.SynchrotronSpectrumFromDist(prtl_dist, esync_bins, 1, 1)
    x_prtls = prtl_dist.EnergyBins().as_array()
    y_prtls = prtl_dist.F().as_array()
    x_sync, y_sync = esync_bins.as_array(), s.as_array()
    assert _check_slope(x_prtls, y_prtls, -p, 2, 800)
    assert _check_slope(x_sync, y_sync, -p / 2 + 3 / 2, 10, 1e4)
    assert _check_slope(x_sync, y_sync, 1 / 3 + 1, 3e-2, 2e-1, atol=0.1)


def test_sync_lin(rg):
    # src/tests/synchrotron.py:38-59
    p = 2.5
    dist = rg.TabulatedDistribution(rg.Linbins(1, 1000, 10000), rg.PlawGenerator(-p, 1, 1000))
    bins = rg.Logbins(0.01, 1e6, 500)
    bins.unit = rg.EnergyUnits.mec2
    s = rg.SynchrotronSpectrumFromDist(dist, bins, 1, 1)
    assert _check_slope(dist.EnergyBins().as_array(), dist.F().as_array(), -p, 2, 800)
    assert _check_slope(bins.as_array(), s.as_array(), -p / 2 + 3 / 2, 10, 1e4)
    assert _check_slope(bins.as_array(), s.as_array(), 1 / 3 + 1, 3e-2, 2e-1, atol=0.1)


def test_sync_prtls(rg, orc):
    # src/tests/synchrotron.py:62-104, plus parity of every output with the oracle
    rng = np.random.default_rng(123)

    def random_plaw(size, xmin, xmax, p, rng=np.random.default_rng(123)):
        return ((xmax ** (p + 1) - xmin ** (p + 1)) * rng.random(size) + xmin ** (p + 1)) ** (1 / (p + 1))

    nprtls = int(1e5)
    rnd1 = 2 * (rng.random(nprtls) - 0.5)
    rnd2 = 2 * np.pi * rng.random(nprtls)
    b1s = np.sqrt(1 - rnd1**2) * np.cos(rnd2)
    b2s = np.sqrt(1 - rnd1**2) * np.sin(rnd2)
    b3s = rnd1
    U1 = random_plaw(size=nprtls, xmin=1, xmax=100, p=-2, rng=rng)
    prtls = rg.Particles_3D("pairs")
    prtls.fromArrays({"U1": U1, "B1": b1s, "B2": b2s, "B3": b3s})
    dist_prtls = prtls.energyDistribution(rg.Logbins(1, 1e3, 100))
    bins_e_syn = rg.Logbins(0.01, 1e5, 200, rg.EnergyUnits.mec2)
    s_p = rg.SynchrotronSpectrum_3D(prtls, bins_e_syn, 1, 1, 1)
    s_d = rg.SynchrotronSpectrumFromDist(dist_prtls, bins_e_syn, 1, 1)
    assert np.abs((s_p.as_array() * 1.3 - s_d.as_array()) / s_d.as_array()).max() < 0.15
    # parity with the oracle on the same float32 inputs
    u = np.zeros((3, nprtls), np.float32)
    u[0] = U1.astype(np.float32)
    b = np.stack([b1s, b2s, b3s]).astype(np.float32)
    _, hd, _ = orc.energy_hist(u, orc.logspace(1, 1e3, 100), True)
    assert np.allclose(dist_prtls.F().as_array(), hd, rtol=1e-6, atol=0)
    ex, f64 = oracle_pair(orc, u, b, orc.logspace(0.01, 1e5, 200))
    assert_spectrum(s_p.as_array(), ex, f64, what="sync_prtls")


def test_sync_unit_check_and_dims(rg):
    prtls = rg.Particles_2D("p")
    prtls.fromArrays({"U1": np.array([10.0]), "B2": np.array([1.0])})
    with pytest.raises(RuntimeError):
        rg.SynchrotronSpectrum_2D(prtls, rg.Logbins(0.01, 1e5, 20), 1, 1, 1)
    s = rg.SynchrotronSpectrum_2D(prtls, rg.Logbins(0.01, 1e5, 20, rg.EnergyUnits.mpc2), 1, 1, 1)
    assert s.extent() == 20 and np.count_nonzero(s.as_array()) > 0
    s1 = rg.SynchrotronSpectrum_1D(rg.Particles_1D("e"), rg.Logbins(0.01, 1e5, 20, "mec2"), 1, 1, 1)
    assert np.count_nonzero(s1.as_array()) == 0


def test_more_than_512_bins(rg, orc):
    prtls = rg.Particles_3D("p")
    rng = np.random.default_rng(0)
    u = rng.normal(0, 30, (3, 5000))
    b = rng.normal(0, 1, (3, 5000))
    prtls.fromArrays({"U1": u[0], "U2": u[1], "U3": u[2], "B1": b[0], "B2": b[1], "B3": b[2]})
    bins = rg.Logbins(1e-2, 1e7, 1200, "mec2")
    s = rg.SynchrotronSpectrum_3D(prtls, bins, 1, 1, 1).as_array()
    ex, f64 = oracle_pair(orc, u.astype(np.float32), b.astype(np.float32), bins.as_array())
    assert_spectrum(s, ex, f64, what="1200 bins")


def test_weights_extension(rg, orc):
    rng = np.random.default_rng(1)
    u = rng.normal(0, 30, (3, 3000))
    b = rng.normal(0, 1, (3, 3000))
    w = rng.uniform(0, 2, 3000)
    p = rg.Particles_3D("w")
    p.fromArrays({"U1": u[0], "U2": u[1], "U3": u[2], "B1": b[0], "B2": b[1], "B3": b[2], "W": w})
    assert p.has_weights()
    bins = rg.Logbins(1e-2, 1e7, 128, "mec2")
    s = rg.SynchrotronSpectrum_3D(p, bins, 1, 1, 1).as_array()
    ex, f64 = oracle_pair(orc, u.astype(np.float32), b.astype(np.float32), bins.as_array(),
                          w=w.astype(np.float32))
    assert_spectrum(s, ex, f64, what="weights")


def test_import_before_torch_uses_one_runtime():
    """Importing ragnar_amd before torch leaves one HIP runtime in the process:
    the drop-in module and torch device tensors both work afterwards."""
    import subprocess
    import sys
    from pathlib import Path

    code = (
        "import numpy as np\n"
        "import ragnar_amd\n"
        "import torch\n"
        "from ragnar_amd import ragnar as rg\n"
        "x = torch.ones(8, device='cuda')\n"
        "assert x.sum().item() == 8.0\n"
        "p = rg.Particles_3D('e')\n"
        "u = np.full(64, 3.0, np.float32)\n"
        "b = np.full(64, 1.0, np.float32)\n"
        "p.fromArrays({'U1': u, 'U2': u * 0, 'U3': u * 0, 'B1': b * 0, 'B2': b, 'B3': b * 0})\n"
        "s = rg.SynchrotronSpectrum_3D(p, rg.Logbins(1e-2, 1e3, 32, 'mec2'), 1.0, 1.0, 1.0)\n"
        "assert np.isfinite(s.as_array()).all() and s.as_array().max() > 0\n"
        "print('ok')\n")
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=300, env={**__import__("os").environ, "RAGNAR_VERBOSE": "0",
                                         "PYTHONPATH": str(root)})
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
