"""Test configuration: import paths and the `gpu` marker.

CPU tests (-m "not gpu") check the oracle against the golden vectors and
closed forms, the host logic and that the C ABI library exports every
declared symbol.  GPU tests (-m gpu) compare the HIP path, called through the
C ABI / Cython package, with the oracle.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "gpu-swimmers_amd", "oracle"):
    p = os.path.join(ROOT, sub) if sub else ROOT
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
