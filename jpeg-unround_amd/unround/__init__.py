"""JPEG decoding that removes compression artifacts within the quantization intervals,
computed on an AMD Instinct GPU (gfx950).

The package mirrors the reference's (python/src/unround/__init__.py) for reconstructing a
JPEG's components given as arrays: unround.solve(image, settings) runs the TV/TGV
primal-dual solve of docs/math.md through the C interface include/unround.h, whose HIP
kernels compute it.

  solve     an Image's components reconstructed (unround.native.solve)
  settings  the method, the model and the solvers' options
  native    the library's functions and types
  jpegio    Image and Component, the arrays a JPEG file holds
  synthetic seeded synthetic pictures, JPEG-quantized, for tests and benchmarks
"""

from unround import jpegio, native, settings
from unround.native import (
    Decoded,
    Record,
    UnroundError,
    UnroundWarning,
    decode,
    solve,
    to_rgb,
    to_ycbcr,
)
from unround.settings import TGV, TV, DataTerm, PdhgOptions, Settings, SubgradientOptions

__all__ = [
    "TGV",
    "TV",
    "DataTerm",
    "Decoded",
    "PdhgOptions",
    "Record",
    "Settings",
    "SubgradientOptions",
    "UnroundError",
    "UnroundWarning",
    "decode",
    "jpegio",
    "native",
    "settings",
    "solve",
    "to_rgb",
    "to_ycbcr",
]
