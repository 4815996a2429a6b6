"""racon_amd: racon's polishing hot path on MI355X (gfx950).

The product is native: libracon_hip.so (HIP kernels behind the C ABI of
include/racon_hip.h) and the racon command line in racon_amd/bin/racon.
This package only exposes the C ABI to Python (ctypes) for tests and the
benchmark driver.
"""
from ._lib import (AlignBatch, Layer, PoaBatch, RACON_BIN, Window, load_library,  # noqa: F401
                   RH_POA_BATCHED, RH_POA_PLAIN, RH_POA_STATIC)

__all__ = ["AlignBatch", "Layer", "PoaBatch", "Window", "load_library", "RACON_BIN",
           "RH_POA_BATCHED", "RH_POA_PLAIN", "RH_POA_STATIC"]
