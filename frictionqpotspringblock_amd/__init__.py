"""MI355X (gfx950) implementation of FrictionQPotSpringBlock's dynamics hot path.

Drop-in for the reference's Python module ``FrictionQPotSpringBlock``: the
submodules ``Line1d``, ``Line2d``, ``Particles`` and ``detail`` expose the same classes,
argument names and defaults (python/main.cpp:228-1006). Every computation runs
in the HIP library ``libfqsb.so`` through its C ABI (include/fqsb.h); there is
no CPU fallback, and importing fails if the native library is not built.

The batched form (many replicas in one resident launch) is in ``batch``.
"""

import os
import sys

_here = os.path.dirname(os.path.abspath(__file__))

# One HIP runtime per process. PyTorch ships its own libamdhip64 (loaded by
# the name libamdhip64.so, with the soname libamdhip64.so.7); loaded first,
# it also satisfies libfqsb.so's dependency on libamdhip64.so.7, so that
# tensors and the library share one runtime and one view of the GPU. Loaded
# after libfqsb.so, torch would bring a second HIP runtime that finds no GPU.
try:
    import torch  # noqa: F401
except ImportError:  # the library itself does not need torch
    pass

try:
    from . import _fqsb
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "frictionqpotspringblock_amd: the native extension is not built "
        f"({e}); run `python -m frictionqpotspringblock_amd.build_ext` first"
    ) from e

from . import _views  # noqa: E402

# u, v, a: writable like the reference's views (item assignments reach the system)
_views.install(_fqsb.System)

Line1d = _fqsb.Line1d
Line2d = _fqsb.Line2d
Particles = _fqsb.Particles
detail = _fqsb.detail
version = _fqsb.version

sys.modules[__name__ + ".Line1d"] = Line1d
sys.modules[__name__ + ".Line2d"] = Line2d
sys.modules[__name__ + ".Particles"] = Particles
sys.modules[__name__ + ".detail"] = detail

LIBRARY_PATH = os.path.join(_here, "libfqsb.so")

__all__ = ["Line1d", "Line2d", "Particles", "detail", "version", "LIBRARY_PATH"]
