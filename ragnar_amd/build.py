"""Build recipe for the gfx950 library and the ``ragnar`` Python module.

Outputs (in-tree, next to this file):
  libragnar_hip.so                 HIP kernels + the C ABI declared in include/ragnar_hip.h
  ragnar<EXT_SUFFIX>               pybind11 module with the reference's Python API

Every HIP translation unit is compiled with -ffp-contract=off (the reference
arithmetic is restated operation by operation; fused multiply-adds are written
explicitly where the kernels want them) and with the packed-FP32 feature
disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["rgn_core.hip", "ftable.hip", "emission.hip", "binned.hip", "cells.hip",
               "histogram.hip", "contraction.hip", "contraction_f16.hip", "ic.hip"]
HOST_SOURCES = ["host/pyragnar.cpp", "host/containers.cpp", "host/h5io.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HDF5_ROOT = Path(os.environ.get("RAGNAR_HDF5_ROOT", "/opt/conda"))

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def _hdf5_available() -> bool:
    return (HDF5_ROOT / "include" / "hdf5.h").exists() and (HDF5_ROOT / "lib" / "libhdf5.so").exists()


def module_path() -> Path:
    return HERE / ("ragnar" + sysconfig.get_config_var("EXT_SUFFIX"))


def library_path() -> Path:
    return HERE / "libragnar_hip.so"


def build(jobs: int | None = None, verbose: bool = False) -> None:
    """Compile every HIP source for gfx950 and the pybind11 module (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)

    def hip_obj(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".o")
        _run([HIPCC, *COMMON, *DEVICE, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    def host_obj(src: str) -> Path:
        import pybind11

        obj = BUILD / (Path(src).stem + ".host.o")
        flags = [*COMMON, f"-I{pybind11.get_include()}",
                 f"-I{sysconfig.get_paths()['include']}", "-D__HIP_PLATFORM_AMD__",
                 "-I/opt/rocm/include", "-fvisibility=hidden"]
        if src.endswith("h5io.cpp") and _hdf5_available():
            flags += [f"-I{HDF5_ROOT / 'include'}", "-DRAGNAR_USE_HDF5=1"]
        _run(["g++", *flags, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        hip_objs = list(ex.map(hip_obj, HIP_SOURCES))
        host_objs = list(ex.map(host_obj, HOST_SOURCES))

    lib = library_path()
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, hip_objs), "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, lib)

    mod = module_path()
    tmp = mod.with_name(mod.name + ".tmp")
    link = ["g++", "-shared", "-fPIC", "-o", str(tmp), *map(str, host_objs),
            f"-L{HERE}", "-lragnar_hip", "-Wl,-rpath,$ORIGIN",
            "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    if _hdf5_available():
        # the system runtime directory goes first so that the HDF5 prefix's own
        # (older) libstdc++ is never picked up through this search path
        link += [f"-L{HDF5_ROOT / 'lib'}", "-lhdf5", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu",
                 f"-Wl,-rpath,{HDF5_ROOT / 'lib'}"]
    _run(link)
    os.replace(tmp, mod)
    if verbose:
        print(f"built {lib.name} and {mod.name}")


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
