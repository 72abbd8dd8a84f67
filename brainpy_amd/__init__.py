"""brainpy_amd -- a gfx950 (MI355X) implementation of BrainPy's event-driven
synaptic hot path behind the reference's operator signatures.

    import brainpy_amd.math as bm
    y = bm.event.csrmv(w, indices, indptr, spikes, shape=(m, n), transpose=True)

Kernels live in ``libbpx.so`` (HIP, C ABI, ``include/bpx.h``); this package
is the host-side mirror of ``brainpy.math.{event,sparse,jitconn}`` plus the
fused LIF/COBA network step (``brainpy_amd.dyn``).  There is no CPU fallback.
"""
from . import _lib
from . import math

__version__ = "0.1.0"
__all__ = ["math", "build"]


def build(verbose=False):
    """Compile libbpx.so for gfx950 in-tree."""
    from .csrc.build import build as _build
    return _build(verbose=verbose)
