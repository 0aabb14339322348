"""A JPEG file's components as arrays: what the reference's reader gives
(python/src/unround/jpegio.py), for unround.solve. Reading files is not part of this
package: make an Image from levels and tables you have (unround.synthetic makes some).
"""

import enum
from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

__all__ = ["ColorSpace", "Component", "Image"]


class ColorSpace(enum.IntEnum):
    """The color spaces of c/jpegio/include/unround/jpegio.h:63-67."""

    GRAYSCALE = 1
    YCBCR = 2
    RGB = 3


@dataclass(frozen=True, slots=True, eq=False)
class Component:
    """A component's levels (rows, columns, 8, 8) int16 in natural order, its quantization
    table (8, 8) or (64,) in natural order, and its sampling factors."""

    coefficients: npt.NDArray[np.int16]
    quant_table: npt.NDArray[np.uint16]
    h_samp_factor: int = 1
    v_samp_factor: int = 1


@dataclass(frozen=True, slots=True, eq=False)
class Image:
    """A picture of height x width samples in a color space, and its components."""

    height: int
    width: int
    color_space: ColorSpace
    components: tuple[Component, ...]
