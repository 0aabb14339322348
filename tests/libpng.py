"""libpng 1.6 through ctypes, as the reference's PNG writer drives it
(c/jpegio/src/png.c:250-276): the system's libpng16 writes the files the library's own
writer must equal byte for byte. LIBPNG is None where no libpng16 is installed."""

from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

_VOID = ctypes.c_void_p
_WRITE = ctypes.CFUNCTYPE(None, _VOID, ctypes.POINTER(ctypes.c_ubyte), ctypes.c_size_t)
_FLUSH = ctypes.CFUNCTYPE(None, _VOID)


def _libpng():
    for name in ("libpng16.so.16", ctypes.util.find_library("png16")):
        if not name:
            continue
        try:
            library = ctypes.CDLL(name)
        except OSError:
            continue
        library.png_get_libpng_ver.restype = ctypes.c_char_p
        library.png_get_libpng_ver.argtypes = (_VOID,)
        library.png_create_write_struct.restype = _VOID
        library.png_create_write_struct.argtypes = (ctypes.c_char_p, _VOID, _VOID, _VOID)
        library.png_create_info_struct.restype = _VOID
        library.png_create_info_struct.argtypes = (_VOID,)
        library.png_set_write_fn.argtypes = (_VOID, _VOID, _WRITE, _FLUSH)
        library.png_set_user_limits.argtypes = (_VOID, ctypes.c_uint32, ctypes.c_uint32)
        library.png_set_compression_level.argtypes = (_VOID, ctypes.c_int)
        library.png_set_IHDR.argtypes = (_VOID, _VOID, ctypes.c_uint32, ctypes.c_uint32) + (ctypes.c_int,) * 5
        library.png_write_info.argtypes = (_VOID, _VOID)
        library.png_write_row.argtypes = (_VOID, ctypes.c_char_p)
        library.png_write_end.argtypes = (_VOID, _VOID)
        library.png_destroy_write_struct.argtypes = (ctypes.POINTER(_VOID), ctypes.POINTER(_VOID))
        return library
    return None


LIBPNG = _libpng()
if LIBPNG is not None and not LIBPNG.png_get_libpng_ver(None).startswith(b"1.6"):
    LIBPNG = None


def libpng_file(levels: np.ndarray, bits: int, compression: int) -> bytes:
    """The file libpng writes for (height, width[, channels]) levels, as png.c drives it:
    colour type from the channels, no interlace, its default filters and zlib settings, the
    level set only when it is not -1, 16-bit rows most significant byte first."""
    height, width = levels.shape[:2]
    channels = 1 if levels.ndim == 2 else levels.shape[2]
    out = bytearray()

    @_WRITE
    def write(_png, data, size):
        out.extend(ctypes.string_at(data, size))

    @_FLUSH
    def flush(_png):
        pass

    png = LIBPNG.png_create_write_struct(LIBPNG.png_get_libpng_ver(None), None, None, None)
    info = LIBPNG.png_create_info_struct(png)
    LIBPNG.png_set_write_fn(png, None, write, flush)
    LIBPNG.png_set_user_limits(png, 0x7FFFFFFF, 0x7FFFFFFF)
    if compression >= 0:
        LIBPNG.png_set_compression_level(png, compression)
    LIBPNG.png_set_IHDR(png, info, width, height, bits, {1: 0, 2: 4, 3: 2, 4: 6}[channels], 0, 0, 0)
    LIBPNG.png_write_info(png, info)
    rows = np.ascontiguousarray(levels.astype(">u2" if bits == 16 else np.uint8)).reshape(height, -1)
    for y in range(height):
        LIBPNG.png_write_row(png, rows[y].tobytes())
    LIBPNG.png_write_end(png, None)
    png_pointer, info_pointer = _VOID(png), _VOID(info)
    LIBPNG.png_destroy_write_struct(ctypes.byref(png_pointer), ctypes.byref(info_pointer))
    return bytes(out)
