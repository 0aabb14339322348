"""The PNG reader of the file layer (unround_jpegio_read_png, c/jpegio/include/unround/
jpegio.h:239-250) without a GPU: files that Pillow writes (its own filters, every colour
type and depth it has), and files written here with every filter and Adam7 interlacing,
each read back to the levels they were made of and, where Pillow reads them, to Pillow's
own reading. Palette files, broken files and pictures beyond the pixel limit are refused
with the layer's statuses; a tRNS chunk is named in the warning; an ICC profile is kept."""

from __future__ import annotations

import ctypes
import io
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from unround import native

OK, ARGUMENT, UNSUPPORTED, LIMIT, PNG = 0, 1, 3, 4, 7


class _PngImage(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("channels", ctypes.c_int32),
                ("bits", ctypes.c_int32), ("samples", ctypes.POINTER(ctypes.c_uint16)),
                ("icc_profile", ctypes.POINTER(ctypes.c_uint8)), ("icc_profile_size", ctypes.c_uint64)]


def library() -> ctypes.CDLL:
    lib = ctypes.CDLL(native.library_path())
    lib.unround_jpegio_read_png.restype = ctypes.c_int32
    lib.unround_jpegio_read_png.argtypes = (ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint64,
                                            ctypes.POINTER(_PngImage), ctypes.c_char_p, ctypes.c_size_t)
    lib.unround_jpegio_png_image_free.argtypes = (ctypes.POINTER(_PngImage),)
    return lib


def read(data: bytes, max_pixels: int = 0) -> tuple[int, np.ndarray | None, int, str, bytes | None]:
    """status, levels (height, width, channels), bits, message, ICC profile."""
    lib = library()
    image = _PngImage()
    message = ctypes.create_string_buffer(256)
    status = lib.unround_jpegio_read_png(data, len(data), max_pixels, ctypes.byref(image), message, 256)
    if status != OK:
        assert not image.samples
        return status, None, 0, message.value.decode(), None
    shape = (image.height, image.width, image.channels)
    levels = np.ctypeslib.as_array(image.samples, shape=(int(np.prod(shape)),)).copy().reshape(shape)
    profile = bytes(image.icc_profile[:image.icc_profile_size]) if image.icc_profile else None
    bits = int(image.bits)
    lib.unround_jpegio_png_image_free(ctypes.byref(image))
    assert not image.samples and image.width == 0
    return status, levels, bits, message.value.decode(), profile


def picture(height: int, width: int, channels: int, top: int, seed: int) -> np.ndarray:
    """Levels with smooth parts and noise, (height, width, channels)."""
    draw = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    smooth = ((x * 3 + y * 5)[..., None] + np.arange(channels) * 17) % (top + 1)
    noise = draw.integers(0, top + 1, size=(height, width, channels))
    mask = draw.random((height, width, 1)) < 0.3
    return np.where(mask, noise, smooth).astype(np.uint16)


# ---------------------------------------------------------------------------------------
# Files that Pillow writes


@pytest.mark.parametrize("mode,channels,bits", [("L", 1, 8), ("LA", 2, 8), ("RGB", 3, 8), ("RGBA", 4, 8),
                                                ("I;16", 1, 16), ("1", 1, 1)])
def test_pillow_files_read_to_their_levels(mode, channels, bits):
    top = (1 << bits) - 1
    levels = picture(37, 53, channels, top, seed=bits * 10 + channels)
    if mode == "I;16":
        made = Image.fromarray(levels[..., 0].astype(np.uint16))
    elif mode == "1":
        made = Image.fromarray(levels[..., 0].astype(bool))
    else:
        made = Image.fromarray(levels.astype(np.uint8) if channels > 1 else levels[..., 0].astype(np.uint8), mode=mode)
    buffer = io.BytesIO()
    made.save(buffer, "PNG")
    status, got, got_bits, message, _ = read(buffer.getvalue())
    assert status == OK and message == ""
    assert got_bits == bits and got.shape == levels.shape
    assert np.array_equal(got, levels)


# ---------------------------------------------------------------------------------------
# Files written here: every filter, Adam7, every depth


def chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def packed_row(row: np.ndarray, bits: int) -> bytes:
    """A row of samples (pixels x channels) as PNG stores it."""
    flat = row.reshape(-1).astype(np.int64)
    if bits == 16:
        return b"".join(struct.pack(">H", int(v)) for v in flat)
    if bits == 8:
        return bytes(int(v) for v in flat)
    out = bytearray((len(flat) * bits + 7) // 8)
    for k, v in enumerate(flat):
        bit = k * bits
        out[bit // 8] |= int(v) << (8 - bits - bit % 8)
    return bytes(out)


def predictor(kind: int, left: int, up: int, corner: int) -> int:
    if kind == 0:
        return 0
    if kind == 1:
        return left
    if kind == 2:
        return up
    if kind == 3:
        return (left + up) // 2
    estimate = left + up - corner
    near = [abs(estimate - left), abs(estimate - up), abs(estimate - corner)]
    return (left, up, corner)[near.index(min(near))]


def filtered(raw: bytes, previous: bytes | None, kind: int, step: int) -> bytes:
    out = bytearray(len(raw))
    for i, value in enumerate(raw):
        left = raw[i - step] if i >= step else 0
        up = previous[i] if previous is not None else 0
        corner = previous[i - step] if previous is not None and i >= step else 0
        out[i] = (value - predictor(kind, left, up, corner)) % 256
    return bytes(out)


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
COLOUR_TYPES = {1: 0, 2: 4, 3: 2, 4: 6}


def png_file(levels: np.ndarray, bits: int, interlaced: bool, extra: list[bytes] = ()) -> bytes:
    height, width, channels = levels.shape
    step = max(1, channels * bits // 8)
    passes = ADAM7 if interlaced else [(0, 0, 1, 1)]
    stream = bytearray()
    row_number = 0
    for x0, y0, dx, dy in passes:
        part = levels[y0::dy, x0::dx]
        if part.size == 0:
            continue
        previous = None
        for row in part:
            raw = packed_row(row, bits)
            kind = row_number % 5
            row_number += 1
            stream.append(kind)
            stream += filtered(raw, previous, kind, step)
            previous = raw
    header = struct.pack(">IIBBBBB", width, height, bits, COLOUR_TYPES[channels], 0, 0, 1 if interlaced else 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", header) + b"".join(extra) +
            chunk(b"IDAT", zlib.compress(bytes(stream), 9)) + chunk(b"IEND", b""))


CASES = [(1, 1), (1, 2), (1, 4), (1, 8), (1, 16), (2, 8), (2, 16), (3, 8), (3, 16), (4, 8), (4, 16)]


@pytest.mark.parametrize("interlaced", [False, True])
@pytest.mark.parametrize("channels,bits", CASES)
def test_every_filter_depth_and_interlacing_reads_back(channels, bits, interlaced):
    levels = picture(13, 19, channels, (1 << bits) - 1, seed=channels * 100 + bits)
    data = png_file(levels, bits, interlaced)
    status, got, got_bits, message, _ = read(data)
    assert status == OK, message
    assert got_bits == bits and np.array_equal(got, levels)
    if bits == 8:
        # Pillow's reading of the same file.
        assert np.array_equal(np.asarray(Image.open(io.BytesIO(data))).reshape(levels.shape), levels)


@pytest.mark.parametrize("height,width", [(1, 1), (1, 9), (9, 1), (2, 3), (7, 5)])
def test_small_interlaced_pictures_have_empty_passes(height, width):
    levels = picture(height, width, 3, 255, seed=height * 10 + width)
    status, got, _, _, _ = read(png_file(levels, 8, True))
    assert status == OK and np.array_equal(got, levels)


# ---------------------------------------------------------------------------------------
# What the reader refuses, and what it notes


def test_refusals_and_warnings():
    levels = picture(6, 7, 3, 255, seed=1)
    good = png_file(levels, 8, False)
    # A palette file (c/jpegio/src/png.c:441-443).
    buffer = io.BytesIO()
    Image.fromarray(levels.astype(np.uint8)).convert("P").save(buffer, "PNG")
    status, _, _, message, _ = read(buffer.getvalue())
    assert status == UNSUPPORTED and "palette" in message
    # Not a PNG file, cut short, a wrong CRC on IDAT, an unknown filter.
    assert read(b"GIF89a" + good[6:])[0] == PNG
    assert read(good[:-30])[0] == PNG
    at = good.index(b"IDAT")
    broken = bytearray(good)
    broken[at + 6] ^= 0xFF
    assert read(bytes(broken))[0] == PNG
    # More pixels than allowed (c/jpegio/src/png.c:446-448).
    status, _, _, message, _ = read(good, max_pixels=41)
    assert status == LIMIT and "pixels" in message
    assert read(good, max_pixels=42)[0] == OK
    # A tRNS chunk is not taken, and said so.
    noted = png_file(levels, 8, False, [chunk(b"tRNS", b"\x00\x01\x00\x02\x00\x03")])
    status, got, _, message, _ = read(noted)
    assert status == OK and np.array_equal(got, levels) and "tRNS" in message
    # A null pointer where the image goes.
    lib = library()
    assert lib.unround_jpegio_read_png(good, len(good), 0, None, None, 0) == ARGUMENT


def test_the_icc_profile_is_kept():
    profile = bytearray(200)
    profile[0:4] = struct.pack(">I", 200)
    profile[36:40] = b"acsp"
    profile[40:200] = bytes(range(160))
    buffer = io.BytesIO()
    Image.fromarray(picture(5, 6, 3, 255, seed=2).astype(np.uint8)).save(buffer, "PNG", icc_profile=bytes(profile))
    status, _, _, message, kept = read(buffer.getvalue())
    assert status == OK and message == "" and kept == bytes(profile)
