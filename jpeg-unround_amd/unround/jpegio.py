"""The file layer (include/unround_jpegio.h) for Python, through ctypes: a JPEG file's
coefficients and metadata, and its component planes, as the reference's module reads them
(python/src/unround/jpegio.py). The results are copied into read-only NumPy arrays, and
the layer's memory is freed before a function returns. The layer lives in the library
itself (unround/_native), on its own JPEG reader.

An Image for unround.solve may also be made from levels and tables one has
(make_component, make_image; unround.synthetic makes some).
"""

from __future__ import annotations

import ctypes
import enum
import functools
from dataclasses import dataclass
from typing import Final

import numpy as np
import numpy.typing as npt

__all__ = [
    "ABI_VERSION",
    "ColorSpace",
    "Component",
    "ErrorKind",
    "Image",
    "JpegioError",
    "Planes",
    "abi_version",
    "check_icc",
    "decode_planes",
    "libjpeg_version",
    "libpng_version",
    "make_component",
    "make_image",
    "read",
]

ABI_VERSION: Final = 3
_MAX_COMPONENTS: Final = 4
_BLOCK_SIZE: Final = 64
_MESSAGE_SIZE: Final = 256


class _Options(ctypes.Structure):
    _fields_ = (("max_pixels", ctypes.c_uint64), ("max_scans", ctypes.c_int32), ("warnings_are_errors", ctypes.c_int32))


class _Component(ctypes.Structure):
    _fields_ = (
        ("id", ctypes.c_int32),
        ("h_samp_factor", ctypes.c_int32),
        ("v_samp_factor", ctypes.c_int32),
        ("quant_table_slot", ctypes.c_int32),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("width_in_blocks", ctypes.c_uint32),
        ("height_in_blocks", ctypes.c_uint32),
        ("quant_table", ctypes.c_uint16 * _BLOCK_SIZE),
        ("coefficients", ctypes.POINTER(ctypes.c_int16)),
    )


class _Image(ctypes.Structure):
    _fields_ = (
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("color_space", ctypes.c_int32),
        ("num_components", ctypes.c_int32),
        ("max_h_samp_factor", ctypes.c_int32),
        ("max_v_samp_factor", ctypes.c_int32),
        ("progressive", ctypes.c_int32),
        ("arithmetic", ctypes.c_int32),
        ("exif_orientation", ctypes.c_int32),
        ("warnings", ctypes.c_int32),
        ("components", _Component * _MAX_COMPONENTS),
        ("icc_profile", ctypes.POINTER(ctypes.c_uint8)),
        ("icc_profile_size", ctypes.c_uint64),
    )


class _Plane(ctypes.Structure):
    _fields_ = (("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("stride", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("samples", ctypes.POINTER(ctypes.c_uint8)))


class _Planes(ctypes.Structure):
    _fields_ = (("num_components", ctypes.c_int32), ("reserved", ctypes.c_int32), ("planes", _Plane * _MAX_COMPONENTS))


_SIZES: Final = ((_Options, 16), (_Component, 168), (_Image, 728), (_Plane, 24), (_Planes, 104))


class ErrorKind(enum.IntEnum):
    """What went wrong, as unround_jpegio_status says it."""

    ARGUMENT = 1
    DECODE = 2
    UNSUPPORTED = 3
    LIMIT = 4
    MEMORY = 5
    ENCODE = 6
    PNG = 7


class ColorSpace(enum.IntEnum):
    """The color space of the components, as the file declares it (jpegio.h:62-67)."""

    GRAYSCALE = 1
    YCBCR = 2
    RGB = 3


class JpegioError(Exception):
    """A file the layer could not read, or would not."""

    def __init__(self, kind: ErrorKind, message: str) -> None:
        super().__init__(f"{message} ({kind.name.lower()})")
        self.kind = kind
        self.message = message


@dataclass(frozen=True, slots=True, eq=False)
class Component:
    """One component of a JPEG file: quant_table (8, 8) in natural order, and coefficients
    (height_in_blocks, width_in_blocks, 8, 8) int16, [by, bx, v, u] the coefficient of
    vertical frequency v and horizontal frequency u; no dummy blocks."""

    id: int
    h_samp_factor: int
    v_samp_factor: int
    quant_table_slot: int
    width: int
    height: int
    quant_table: npt.NDArray[np.uint16]
    coefficients: npt.NDArray[np.int16]

    @property
    def width_in_blocks(self) -> int:
        return int(self.coefficients.shape[1])

    @property
    def height_in_blocks(self) -> int:
        return int(self.coefficients.shape[0])


@dataclass(frozen=True, slots=True, eq=False)
class Image:
    """What read() returns (python/src/unround/jpegio.py:190-209)."""

    width: int
    height: int
    color_space: ColorSpace
    max_h_samp_factor: int
    max_v_samp_factor: int
    progressive: bool
    arithmetic: bool
    exif_orientation: int
    warnings: int
    warning: str
    components: tuple[Component, ...]
    icc_profile: bytes | None


@dataclass(frozen=True, slots=True, eq=False)
class Planes:
    """What decode_planes() returns: one uint8 array per component, the accurate integer
    inverse DCT of its blocks, level-shifted and clamped, padded to whole blocks."""

    planes: tuple[npt.NDArray[np.uint8], ...]
    warning: str


def make_component(levels: npt.ArrayLike, quant_table: npt.ArrayLike, h_samp_factor: int = 1,
                   v_samp_factor: int = 1, *, id: int = 1, quant_table_slot: int = 0) -> Component:  # noqa: A002
    """A component of levels (rows, columns, 8, 8) and its table (64 or 8 x 8, natural
    order), its sample sizes those of its blocks."""
    coefficients = np.ascontiguousarray(levels, dtype=np.int16)
    table = np.ascontiguousarray(quant_table, dtype=np.uint16).reshape(8, 8)
    rows, columns = (int(coefficients.shape[0]), int(coefficients.shape[1])) if coefficients.ndim == 4 else (0, 0)
    return Component(id, int(h_samp_factor), int(v_samp_factor), quant_table_slot, 8 * columns, 8 * rows, table,
                     coefficients)


def make_image(height: int, width: int, color_space: ColorSpace, components: tuple[Component, ...]) -> Image:
    """An Image of components one has, for unround.solve."""
    return Image(width=int(width), height=int(height), color_space=ColorSpace(int(color_space)),
                 max_h_samp_factor=max((c.h_samp_factor for c in components), default=1),
                 max_v_samp_factor=max((c.v_samp_factor for c in components), default=1), progressive=False,
                 arithmetic=False, exif_orientation=0, warnings=0, warning="", components=tuple(components),
                 icc_profile=None)


@functools.cache
def _library() -> ctypes.CDLL:
    from unround import native

    for structure, size in _SIZES:
        if ctypes.sizeof(structure) != size:
            raise RuntimeError(f"{structure.__name__} is {ctypes.sizeof(structure)} bytes, the layer's {size}")
    library = native._lib()
    message = (ctypes.c_char_p, ctypes.c_size_t)
    library.unround_jpegio_abi_version.restype = ctypes.c_int32
    library.unround_jpegio_libjpeg_version.restype = ctypes.c_char_p
    library.unround_jpegio_libpng_version.restype = ctypes.c_char_p
    library.unround_jpegio_check_icc.restype = ctypes.c_int32
    library.unround_jpegio_check_icc.argtypes = (ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int32, *message)
    for name, structure in (("read", _Image), ("decode_planes", _Planes)):
        function = getattr(library, "unround_jpegio_" + name)
        function.restype = ctypes.c_int32
        function.argtypes = (ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_Options), ctypes.POINTER(structure),
                             *message)
    library.unround_jpegio_image_free.argtypes = (ctypes.POINTER(_Image),)
    library.unround_jpegio_planes_free.argtypes = (ctypes.POINTER(_Planes),)
    return library


def abi_version() -> int:
    return int(_library().unround_jpegio_abi_version())


def libjpeg_version() -> str:
    return _library().unround_jpegio_libjpeg_version().decode("ascii")


def libpng_version() -> str:
    return _library().unround_jpegio_libpng_version().decode("ascii")


def check_icc(profile: bytes, channels: int) -> str | None:
    """Why an ICC profile does not go with a picture of that many channels, or None."""
    reason = ctypes.create_string_buffer(_MESSAGE_SIZE)
    taken = _library().unround_jpegio_check_icc(profile, len(profile), channels, reason, _MESSAGE_SIZE)
    return None if taken == 1 else reason.value.decode("utf-8", "replace")


def _options(max_pixels: int, max_scans: int, warnings_are_errors: bool) -> _Options:
    if max_pixels < 0 or max_pixels >= 1 << 64:
        raise ValueError(f"max_pixels must fit an unsigned 64-bit count; got {max_pixels}")
    if max_scans < 0 or max_scans >= 1 << 31:
        raise ValueError(f"max_scans must fit a non-negative 32-bit int; got {max_scans}")
    return _Options(max_pixels, max_scans, 1 if warnings_are_errors else 0)


def _raise_for(status: int, message: str) -> None:
    if status == 0:
        return
    if status in ErrorKind.__members__.values():
        raise JpegioError(ErrorKind(status), message)
    raise RuntimeError(f"the layer returned status {status}: {message}")


def _read_only(array: np.ndarray) -> np.ndarray:
    array.flags.writeable = False
    return array


def read(data: bytes | bytearray | memoryview, *, max_pixels: int = 0, max_scans: int = 0,
         warnings_are_errors: bool = False) -> Image:
    """The coefficients and metadata of a JPEG file in memory (python/src/unround/jpegio.py:
    386-425); JpegioError where it cannot be read."""
    library = _library()
    buffer = bytes(data)
    options = _options(max_pixels, max_scans, warnings_are_errors)
    raw = _Image()
    message = ctypes.create_string_buffer(_MESSAGE_SIZE)
    status = library.unround_jpegio_read(buffer, len(buffer), ctypes.byref(options), ctypes.byref(raw), message,
                                         _MESSAGE_SIZE)
    text = message.value.decode("utf-8", "replace")
    _raise_for(int(status), text)
    try:
        components = []
        for index in range(int(raw.num_components)):
            c = raw.components[index]
            down, across = int(c.height_in_blocks), int(c.width_in_blocks)
            count = down * across * _BLOCK_SIZE
            values = np.array(np.ctypeslib.as_array(c.coefficients, shape=(count,)), dtype=np.int16, copy=True)
            table = np.array(np.ctypeslib.as_array(c.quant_table), dtype=np.uint16, copy=True)
            components.append(Component(int(c.id), int(c.h_samp_factor), int(c.v_samp_factor), int(c.quant_table_slot),
                                        int(c.width), int(c.height), _read_only(table.reshape(8, 8)),
                                        _read_only(values.reshape(down, across, 8, 8))))
        profile = ctypes.string_at(raw.icc_profile, int(raw.icc_profile_size)) if raw.icc_profile else None
        return Image(width=int(raw.width), height=int(raw.height), color_space=ColorSpace(int(raw.color_space)),
                     max_h_samp_factor=int(raw.max_h_samp_factor), max_v_samp_factor=int(raw.max_v_samp_factor),
                     progressive=bool(raw.progressive), arithmetic=bool(raw.arithmetic),
                     exif_orientation=int(raw.exif_orientation), warnings=int(raw.warnings), warning=text,
                     components=tuple(components), icc_profile=profile)
    finally:
        library.unround_jpegio_image_free(ctypes.byref(raw))


def decode_planes(data: bytes | bytearray | memoryview, *, max_pixels: int = 0, max_scans: int = 0,
                  warnings_are_errors: bool = False) -> Planes:
    """The component planes of a JPEG file in memory, with read()'s limits."""
    library = _library()
    buffer = bytes(data)
    options = _options(max_pixels, max_scans, warnings_are_errors)
    raw = _Planes()
    message = ctypes.create_string_buffer(_MESSAGE_SIZE)
    status = library.unround_jpegio_decode_planes(buffer, len(buffer), ctypes.byref(options), ctypes.byref(raw),
                                                  message, _MESSAGE_SIZE)
    text = message.value.decode("utf-8", "replace")
    _raise_for(int(status), text)
    try:
        planes = []
        for index in range(int(raw.num_components)):
            plane = raw.planes[index]
            height, width, stride = int(plane.height), int(plane.width), int(plane.stride)
            samples = np.ctypeslib.as_array(plane.samples, shape=(height * stride,)).reshape(height, stride)
            planes.append(_read_only(np.array(samples[:, :width], dtype=np.uint8, copy=True)))
        return Planes(planes=tuple(planes), warning=text)
    finally:
        library.unround_jpegio_planes_free(ctypes.byref(raw))
