"""Swimmer arrays (reference: lbswim/swimmers.pyx, lbswim/SwimmerArray.pxd)."""
from ._core import CommonParams, SwimmerArray as Array  # noqa: F401
