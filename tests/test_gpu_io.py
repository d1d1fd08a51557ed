"""HDF5 I/O and the Tristan-v2 particle reader (plugins/tristan-v2.cpp:51-258,
io/h5.cpp:16-118).  The reference never tests these (its CI builds without
HDF5); here a synthetic Tristan-v2 dump is written with H5write1DArray_f and
read back, with the start/size/stride semantics and error cases of the
reference, and a file larger than the 16 MiB staging buffer exercises the
pinned double-buffered path into device memory.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QUANTITIES = ["x", "y", "z", "u", "v", "w", "ex", "ey", "ez", "bx", "by", "bz"]


@pytest.fixture(scope="module")
def rg():
    from ragnar_amd import ragnar as rg

    if not rg.HAS_HDF5:
        pytest.skip("module built without HDF5")
    return rg


def write_dump(rg, root, step, sp, n, seed=0):
    d = root / "output" / "prtl"
    d.mkdir(parents=True, exist_ok=True)
    fname = str(d / f"prtl.tot.{step:05d}")
    rng = np.random.default_rng(seed)
    data = {}
    for q in QUANTITIES:
        arr = rng.normal(0, 10, n).astype(np.float32)
        data[q] = arr
        rg.H5write1DArray_f(fname, f"{q}_{sp}", rg.Array1D_f(arr))
    return fname, data


def test_h5_array_roundtrip(rg, tmp_path):
    f = str(tmp_path / "arrays.h5")
    a_i = np.arange(1000, dtype=np.int32)
    a_f = np.linspace(-1, 1, 777).astype(np.float32)
    a_d = np.logspace(-3, 3, 555)
    rg.H5write1DArray_i(f, "ints", rg.Array1D_i(a_i))
    rg.H5write1DArray_f(f, "floats", rg.Array1D_f(a_f))
    rg.H5write1DArray_d(f, "doubles", rg.Array1D_d(a_d))
    assert np.array_equal(rg.H5read1DArray_i(f, "ints").as_array(), a_i)
    assert np.array_equal(rg.H5read1DArray_f(f, "floats").as_array(), a_f)
    assert np.array_equal(rg.H5read1DArray_d(f, "doubles").as_array(), a_d)
    # size / stride (io/h5.cpp:27-40): stride 2 needs an explicit size >= dims / 2
    got = rg.H5read1DArray_i(f, "ints", 500, 2).as_array()
    assert np.array_equal(got, a_i[::2][:500])
    with pytest.raises(RuntimeError):
        rg.H5read1DArray_i(f, "ints", 10, 2)  # dims / stride > size
    with pytest.raises(RuntimeError):
        rg.H5read1DArray_i(f, "ints", 0, 0)
    with pytest.raises(RuntimeError):
        rg.H5read1DArray_f(f, "missing")
    with pytest.raises(RuntimeError):
        rg.H5write1DArray_f(f, "floats", rg.Array1D_f(a_f))  # exists


def test_tristan_roundtrip(rg, tmp_path):
    n = 3001
    _, data = write_dump(rg, tmp_path, 12, 1, n)
    tr = rg.TristanV2_3D()
    assert tr.label() == "Tristan V2"
    with pytest.raises(RuntimeError):
        tr.getStep()
    tr.setPath(str(tmp_path))
    tr.setStep(12)
    assert tr.getPath() == str(tmp_path) and tr.getStep() == 12
    p = tr.readParticles("e", 1)
    assert p.nactive() == n and p.label() == "e"
    for d, q in enumerate("xyz"):
        assert np.array_equal(p.X(d + 1).as_array(), data[q])
    for d, (qu, qe, qb) in enumerate(zip("uvw", ["ex", "ey", "ez"], ["bx", "by", "bz"])):
        assert np.array_equal(p.U(d + 1).as_array(), data[qu])
        assert np.array_equal(p.E(d + 1).as_array(), data[qe])
        assert np.array_equal(p.B(d + 1).as_array(), data[qb])
    # stride: size is ntotal / stride
    p3 = tr.readParticles("s", 1, stride=3)
    assert p3.nactive() == n // 3
    assert np.array_equal(p3.U(1).as_array(), data["u"][::3][: n // 3])
    # start + size
    ps = tr.readParticles("w", 1, start=10, size=100)
    assert np.array_equal(ps.B(3).as_array(), data["bz"][10:110])
    # ignore_coordinates
    pi = tr.readParticles("i", 1, ignore_coordinates=True)
    with pytest.raises(RuntimeError):
        pi.X(1)
    assert np.array_equal(pi.U(2).as_array(), data["v"])
    # errors (tristan-v2.cpp:102-107, 126-128)
    with pytest.raises(RuntimeError):
        tr.readParticles("e", 1, stride=0)
    with pytest.raises(RuntimeError):
        tr.readParticles("e", 1, size=5, stride=2)
    with pytest.raises(RuntimeError):
        tr.readParticles("e", 1, start=n - 5, size=10)
    with pytest.raises(RuntimeError):
        tr.readParticles("e", 2)  # no such species


def test_tristan_large_file_and_physics(rg, tmp_path):
    # 6M particles: 24 MB per quantity, two 16 MiB staging chunks per dataset
    n = 6_000_000
    _, data = write_dump(rg, tmp_path, 3, 2, n, seed=5)
    tr = rg.TristanV2_3D()
    tr.setPath(str(tmp_path))
    tr.setStep(3)
    p = tr.readParticles("big", 2, ignore_coordinates=True)
    assert np.array_equal(p.U(3).as_array(), data["w"])
    assert np.array_equal(p.B(1).as_array(), data["bx"])
    # the same particles through fromArrays give the same spectrum
    q = rg.Particles_3D("same")
    q.fromArrays({"U1": data["u"], "U2": data["v"], "U3": data["w"], "E1": data["ex"],
                  "E2": data["ey"], "E3": data["ez"], "B1": data["bx"], "B2": data["by"],
                  "B3": data["bz"]})
    bins = rg.Logbins(1e-2, 1e7, 128, "mec2")
    s1 = rg.SynchrotronSpectrum_3D(p, bins, 1, 1, 1).as_array()
    s2 = rg.SynchrotronSpectrum_3D(q, bins, 1, 1, 1).as_array()
    assert np.array_equal(s1, s2)
