"""lbswim for MI355X: the coupled D3Q15 lattice-Boltzmann / swimmer / tracer
step on hand-written gfx950 HIP kernels, with the reference package's module
layout (lbswim.lb, lbswim.shared, lbswim.swimmers, lbswim.tracers,
lbswim.System).

Importing fails loudly when the native extension (lbswim/_core*.so, linked to
lib/liblbswim_amd.so) has not been built: there is no CPU fallback.
"""
try:
    from . import _core  # noqa: F401
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "lbswim native extension is not built (run gpu-swimmers_amd/build_native.py): %s" % exc
    ) from exc

from ._core import LbswimError, synchronize  # noqa: E402,F401

__all__ = ["lb", "shared", "swimmers", "tracers", "System", "LbswimError", "synchronize"]
