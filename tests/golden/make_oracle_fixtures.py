"""Regenerate tests/golden/coupled_small.npz from the oracle.

This fixture guards the oracle against silent change; it is pinned itself by
tests/test_oracle.py.  Usage: python tests/golden/make_oracle_fixtures.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import lbswim_oracle as O  # noqa: E402
from golden_cases import coupled_small  # noqa: E402

if __name__ == "__main__":
    out = coupled_small(O)
    np.savez_compressed(os.path.join(HERE, "coupled_small.npz"), **out)
    print("wrote coupled_small.npz", {k: v.shape for k, v in out.items()})
