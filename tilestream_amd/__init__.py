"""TileStream for AMD Instinct MI355X (gfx950).

A lossless tiled codec for uint16 and bit-exact float32 images (.tsc 1.0),
with GPU decode (whole image, region, fused calibration), a canonical GPU
encoder and standalone calibration, all in hand-written HIP kernels reached
through the C-ABI of include/tilestream_c.h.  This package is the Python
binding of that C-ABI; the native library is built in-tree with
`python -m tilestream_amd.build`.
"""
from .api import *  # noqa: F401,F403
from .api import Status, PixelType, TileMode, Operation  # noqa: F401

__version__ = "0.1.0"
