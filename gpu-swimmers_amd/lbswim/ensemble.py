"""Batched ensembles of independent simulations (no reference counterpart)."""
from ._core import Ensemble  # noqa: F401

ORDERS = {0: "member-major", 1: "round-robin", 2: "largest-first"}
