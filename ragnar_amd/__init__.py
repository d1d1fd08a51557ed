"""MI355X-native (gfx950) radiation hot path of ragnar.

``ragnar_amd.ragnar`` is a drop-in for the reference's ``ragnar`` module
(``from ragnar_amd import ragnar as rg``); ``ragnar_amd.torch_ops`` exposes the
same kernels on torch device tensors, including the differentiable
``SynchrotronSpectrumFn``.  Both sit on libragnar_hip.so, whose C ABI is
declared in include/ragnar_hip.h.  There is no CPU execution path: without the
compiled library the import fails.
"""
from __future__ import annotations

from pathlib import Path

_HERE = Path(__file__).resolve().parent

# One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64 and its
# libraries ask for it by the unversioned name, which does not match the
# versioned name this package links; so if this package loaded the system
# runtime first, importing torch afterwards would load a second runtime and
# the device would be unusable from ours.  Loading torch first makes both use
# torch's runtime.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the ragnar module
    pass

try:
    from . import ragnar  # noqa: F401  (pybind11 module, links libragnar_hip.so)
except ImportError as exc:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "ragnar_amd: the compiled module is missing or stale; build it with "
        "`python ragnar_amd/build.py` (or __graft_entry__.build())") from exc

__all__ = ["ragnar"]
