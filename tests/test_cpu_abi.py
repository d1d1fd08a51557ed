"""CPU checks of the boundary: the C-ABI library loads and exports every symbol
the header declares, the Python module imports with the reference's names, and
host-only behaviour (argument errors, generator checks) matches the reference.
No compute call is made here (no GPU in this environment)."""
import ctypes

import numpy as np
import pytest


def test_library_exports_every_declared_symbol():
    from ragnar_amd import _capi

    lib = _capi.lib()
    declared = _capi.declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # every declared symbol also has a ctypes signature
    assert not [s for s in declared if s not in _capi._SIGNATURES]
    assert lib.rgn_version().decode().startswith("ragnar_hip")


def test_module_surface_matches_reference_names():
    from ragnar_amd import ragnar as rg

    # pyinterface.cpp:26-91 and the classes/functions it registers
    names = ["Initialize", "Finalize", "EnergyUnits", "Linspace", "Logspace", "TabulatedFunction",
             "TabulatedFunction_log", "Array1D_i", "Array1D_f", "Array1D_d", "Bins", "Linbins",
             "Logbins", "Particles_1D", "Particles_2D", "Particles_3D", "PlawGenerator",
             "BrokenPlawGenerator", "DeltaGenerator", "TabulatedDistribution", "Ffunc_integrand",
             "SynchrotronSpectrumFromDist", "SynchrotronSpectrum_1D", "SynchrotronSpectrum_2D",
             "SynchrotronSpectrum_3D", "ICSpectrum"]
    missing = [n for n in names if not hasattr(rg, n)]
    assert not missing, missing
    if rg.HAS_HDF5:
        for n in ["H5read1DArray_i", "H5read1DArray_f", "H5read1DArray_d", "H5write1DArray_i",
                  "H5write1DArray_f", "H5write1DArray_d", "TristanV2_1D", "TristanV2_2D",
                  "TristanV2_3D"]:
            assert hasattr(rg, n), n
    assert rg.EnergyUnits.mec2 == "mec2" and rg.EnergyUnits.mpc2 == "mpc2"
    assert rg.EnergyUnits.eV == "eV" and rg.EnergyUnits.MeV == "MeV" and rg.EnergyUnits.GeV == "GeV"


def test_host_only_behaviour():
    from ragnar_amd import ragnar as rg

    # distributions.cpp:42-58: constructor checks raise RuntimeError
    with pytest.raises(RuntimeError, match="emin < 0.0"):
        rg.PlawGenerator(-2.0, -1.0, 5.0)
    with pytest.raises(RuntimeError, match="emin > emax"):
        rg.PlawGenerator(-2.0, 5.0, 1.0)
    with pytest.raises(RuntimeError, match="normalization diverges"):
        rg.PlawGenerator(-2.0)
    with pytest.raises(RuntimeError, match="normalization diverges"):
        rg.PlawGenerator(-0.5, 1.0)
    with pytest.raises(RuntimeError, match="e_break <= 0.0"):
        rg.BrokenPlawGenerator(0.0, -1.0, -2.0)
    g = rg.PlawGenerator(-2.0, 1.0, 100.0)
    assert g.p == -2.0 and g.emin == 1.0 and g.emax == 100.0
    d = rg.DeltaGenerator(0.5, 0.1)
    assert np.isclose(d.energy0, 0.5) and np.isclose(d.denergy, 0.1)
    b = rg.Bins("mec2")
    assert b.unit == "mec2" and not b.log_spaced and b.extent() == 0
    p = rg.Particles_3D("e")
    assert not p.is_allocated() and p.label() == "e" and len(p) == 0
    assert repr(p) == "Particles<3D> (e) : not allocated"
    with pytest.raises(RuntimeError, match="Particles not allocated"):
        p.setNactive(3)


def test_oracle_is_not_imported_by_the_product():
    import ast
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent / "ragnar_amd"
    for f in root.rglob("*.py"):
        tree = ast.parse(f.read_text())
        for node in ast.walk(tree):
            if isinstance(node, (ast.Import, ast.ImportFrom)):
                mod = getattr(node, "module", None) or ""
                names = [a.name for a in node.names]
                assert not mod.startswith("oracle") and not any(n.startswith("oracle") for n in names), f
    for f in list(root.rglob("*.cpp")) + list(root.rglob("*.hip")) + list(root.rglob("*.hpp")):
        assert "oracle" not in f.read_text().lower(), f


def test_product_fails_loudly_without_the_hip_library(tmp_path):
    """A copy of the package without the compiled library and module: the
    package import, the ctypes binding and the torch wrappers all raise
    ImportError.  There is no CPU fallback to land on."""
    import os
    import shutil
    import subprocess
    import sys
    from pathlib import Path

    src = Path(__file__).resolve().parent.parent / "ragnar_amd"
    dst = tmp_path / "ragnar_amd"
    dst.mkdir()
    for f in src.glob("*.py"):
        shutil.copy(f, dst / f.name)
    probe = (
        "import importlib, importlib.util, sys\n"
        "try:\n"
        "    import ragnar_amd\n"
        "    print('package imported')\n"
        "except ImportError:\n"
        "    print('package ImportError')\n"
        # the binding module on its own, outside the package's __init__
        "spec = importlib.util.spec_from_file_location('capi_probe', sys.argv[1])\n"
        "capi = importlib.util.module_from_spec(spec)\n"
        "spec.loader.exec_module(capi)\n"
        "try:\n"
        "    capi.lib()\n"
        "    print('library loaded')\n"
        "except ImportError:\n"
        "    print('library ImportError')\n"
    )
    env = {**os.environ, "PYTHONPATH": str(tmp_path)}
    res = subprocess.run([sys.executable, "-c", probe, str(dst / "_capi.py")], cwd=tmp_path, env=env,
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert "package ImportError" in res.stdout, res.stdout
    assert "library ImportError" in res.stdout, res.stdout
