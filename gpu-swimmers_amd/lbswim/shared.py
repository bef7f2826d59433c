"""Host/device field wrappers (reference: lbswim/shared.pyx, lbswim/_shared.pxd)."""
from ._core import Array, PrngArray  # noqa: F401
