"""Lattice module (reference: lbswim/lb.pyx, lbswim/_lb.pxd)."""
from ._core import LatticeAddressing, LBParams, Lattice, SharedLatticeArray  # noqa: F401

DQ_d = 3
DQ_q = 15
