"""Tracer arrays (reference: lbswim/tracers.pyx, lbswim/TracerArray.pxd)."""
from ._core import TracerArray as Array  # noqa: F401
