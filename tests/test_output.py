"""The pictures the library gives out, without a GPU (docs/math.md 10; rust/unround/src/
output.rs, tiff.rs, orientation.rs; c/jpegio/src/png.c): rounding and the ordered dither
exactly, against the levels the rationals fix; the blue-noise matrix against the digest of
the reference's (tests/golden/blue_noise.json), its ranks read back through the dither
itself; PNM bytes; PNG files read back by Pillow and by the library's reader; the TIFF's
tags and samples bit for bit; the ICC check; and the eight EXIF orientations against
numpy's turns."""

from __future__ import annotations

import ctypes
import hashlib
import io
import json
import math
import struct
import warnings
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest
from PIL import Image

import helpers
from unround import native

GOLDEN = Path(__file__).resolve().parent / "golden"


def levels_of(data: bytes, height: int, width: int, channels: int, sixteen: bool) -> np.ndarray:
    """The samples of a PNM file the library wrote."""
    header = f"{'P5' if channels == 1 else 'P6'}\n{width} {height}\n{65535 if sixteen else 255}\n".encode()
    assert data.startswith(header)
    body = np.frombuffer(data[len(header):], dtype=">u2" if sixteen else np.uint8)
    return body.astype(np.int64).reshape(height, width, channels)


def quantized(samples: np.ndarray, sixteen: bool = False, dither: str = "none",
              together: bool = False) -> np.ndarray:
    picture = np.asarray(samples, dtype=np.float64)
    if picture.ndim == 2:
        picture = picture[..., None]
    height, width, channels = picture.shape
    data = native.pnm_bytes(picture if channels > 1 else picture[..., 0], sixteen=sixteen, dither=dither,
                            dither_channels="together" if together else None)
    return levels_of(data, height, width, channels, sixteen)


def exact_round(x: float, scale: int, top: int) -> int:
    """The nearest level to scale x, halves away from 0, in 0..top, from the rationals."""
    if x < 0:
        return 0
    return min(math.floor(Fraction(x) * scale + Fraction(1, 2)), top)


def exact_dither(x: float, scale: int, rank: int, top: int) -> int:
    """floor(y) + [y - floor(y) >= 1 - (2R + 1) / 8192], y = scale x exactly, in 0..top."""
    if x < 0:
        return 0
    y = Fraction(x) * scale
    whole = math.floor(y)
    return min(whole + (y - whole >= 1 - Fraction(2 * rank + 1, 8192)), top)


# ---------------------------------------------------------------------------------------
# Rounding (rust/unround/src/output.rs:401-410, 545-571)


def test_rounding_takes_halves_away_from_zero_and_clamps():
    samples = np.array([[0.49999999999999994, 0.5, 1.5, 2.5, -0.4, -7.0, 254.5, 255.4, 1e300]])
    assert quantized(samples)[0, :, 0].tolist() == [0, 1, 2, 3, 0, 0, 255, 255, 255]
    assert quantized(np.array([[[255.0, 1.0, 0.0]]]), sixteen=True)[0, 0].tolist() == [65535, 257, 0]
    for bad in (math.nan, math.inf):
        with pytest.raises(native.UnroundError) as caught:
            native.pnm_bytes(np.array([[bad]]))
        assert caught.value.status == native.Status.OPTIONS


def test_sixteen_bits_round_257_times_the_sample_exactly():
    samples = []
    for whole in range(0, 65535, 61):
        near = (whole + 0.5) / 257.0
        samples += [math.nextafter(near, -math.inf), near, math.nextafter(near, math.inf)]
    made = quantized(np.array([samples]), sixteen=True)[0, :, 0]
    assert made.tolist() == [exact_round(x, 257, 65535) for x in samples]
    for x in (0.5, math.nextafter(0.5, 0), 254.5, math.nextafter(254.5, 0)):
        assert quantized(np.array([[x]]))[0, 0, 0] == exact_round(x, 1, 255)


# ---------------------------------------------------------------------------------------
# The blue-noise matrix and the ordered dither


def ranks_read_back() -> np.ndarray:
    """The ranks of the library's matrix, 64 x 64, read through the dither of a grey picture:
    a sample of fraction 1 - (2m + 1) / 8192 goes up where the rank is at least m, so that a
    binary search of 12 pictures finds every rank."""
    low = np.zeros((64, 64), dtype=np.int64)
    high = np.full((64, 64), 4095, dtype=np.int64)
    while np.any(low < high):
        middle = (low + high + 1) // 2
        picture = 1.0 - (2.0 * middle + 1.0) / 8192.0
        up = quantized(picture, dither="blue-noise")[..., 0] == 1
        low = np.where(up, middle, low)
        high = np.where(up, high, middle - 1)
    return low


@pytest.fixture(scope="module")
def ranks() -> np.ndarray:
    return ranks_read_back()


def test_the_matrix_is_the_reference_s(ranks):
    golden = json.loads((GOLDEN / "blue_noise.json").read_text())
    flat = ranks.reshape(-1)
    assert sorted(flat.tolist()) == list(range(4096))
    assert [int(np.flatnonzero(flat == rank)[0]) for rank in range(8)] == golden["places_of_ranks_0_to_7"]
    assert hashlib.sha256(flat.astype("<u2").tobytes()).hexdigest() == golden["sha256_u16le"]


def test_a_dithered_level_is_the_exact_one(ranks):
    # rust/unround/src/output.rs:425-457: the thresholds' own ends and their neighbours, and
    # others; each sample of a flat 64 x 64 picture against its rank's exact level.
    values = [0.0, 1.0, 0.5, 254.99999999999997, 255.0, 300.0, 0.37, 17.9]
    for rank in (0, 1, 2047, 2048, 4094, 4095):
        end = 1.0 - (2.0 * rank + 1.0) / 8192.0
        for whole in (0.0, 7.0, 254.0):
            at = whole + end
            values += [at, math.nextafter(at, -math.inf), math.nextafter(at, math.inf)]
    flat_ranks = ranks.reshape(-1)
    for value in values:
        eight = quantized(np.full((64, 64), value), dither="blue-noise")[..., 0].reshape(-1)
        expected = [exact_dither(value, 1, int(r), 255) for r in flat_ranks]
        assert eight.tolist() == expected, value
    for value in (-1e-300, -0.3, -1e9):
        assert np.all(quantized(np.full((64, 64), value), dither="blue-noise") == 0)


def test_sixteen_bits_dither_257_times_the_sample(ranks):
    draw = np.random.default_rng(3)
    picture = draw.random((64, 64)) * 255.0
    picture[0, :8] = [0.0, 255.0, 1.0 / 257.0, 128.5, -3.0, 256.0, 65534.6 / 257.0, 1e-6]
    made = quantized(picture, sixteen=True, dither="blue-noise")[..., 0]
    for i in range(0, 64, 3):
        for j in range(64):
            assert made[i, j] == exact_dither(float(picture[i, j]), 257, int(ranks[i, j]), 65535), (i, j)


def test_over_the_matrix_a_level_keeps_its_mean():
    # rust/unround/src/output.rs:460-480: a fraction e / 8192 goes up at ceil(e / 2) places.
    for whole, eighths in ((3.0, 0), (3.0, 1), (100.0, 8191), (100.0, 4096), (0.0, 5), (200.0, 6000)):
        y = whole + eighths / 8192.0
        up = int(np.sum(quantized(np.full((64, 64), y), dither="blue-noise") > whole))
        assert up == (eighths + 1) // 2, y


def test_the_channels_take_the_matrix_shifted_apart_or_alike(ranks):
    # rust/unround/src/output.rs:483-505, the shifts (0, 0), (0, 32), (32, 0), (32, 32).
    picture = np.full((70, 70, 3), 9.25)
    apart = quantized(picture, dither="blue-noise")
    together = quantized(picture, dither="blue-noise", together=True)
    shifts = [(0, 0), (0, 32), (32, 0), (32, 32)]
    for c, (down, across) in enumerate(shifts[:3]):
        rows = (np.arange(70)[:, None] + down) % 64
        columns = (np.arange(70)[None, :] + across) % 64
        threshold = (2.0 * ranks[rows, columns] + 1.0) / 8192.0
        assert np.array_equal(apart[..., c], 9 + (1.0 - threshold <= 0.25))
        base = (2.0 * ranks[np.arange(70)[:, None] % 64, np.arange(70)[None, :] % 64] + 1.0) / 8192.0
        assert np.array_equal(together[..., c], 9 + (1.0 - base <= 0.25))


# ---------------------------------------------------------------------------------------
# Files


def test_a_pnm_file_has_its_header_and_samples():
    # rust/unround/src/output.rs:574-591
    assert native.pnm_bytes(np.array([[0.0, 255.0, 128.4]]), dither="none") == b"P5\n3 1\n255\n\x00\xff\x80"
    assert (native.pnm_bytes(np.array([[[1.0, 2.0, 3.0]]]), sixteen=True, dither="none") ==
            b"P6\n1 1\n65535\n\x01\x01\x02\x02\x03\x03")
    assert native.pnm_bytes(np.array([[0.0, 255.0, 128.0]])) == b"P5\n3 1\n255\n\x00\xff\x80"
    with pytest.raises(native.UnroundError):
        native.pnm_bytes(np.zeros((2, 2, 2)))


def profile(space: bytes = b"RGB ", size: int = 200) -> bytes:
    """A small ICC profile that passes the check: its size, 'acsp', a data colour space, XYZ."""
    made = bytearray(size)
    made[0:4] = struct.pack(">I", size)
    made[12:16] = b"mntr"
    made[16:20] = space
    made[20:24] = b"XYZ "
    made[36:40] = b"acsp"
    made[128:132] = struct.pack(">I", 1)
    made[132:144] = b"desc" + struct.pack(">II", 144, 8)
    return bytes(made)


@pytest.mark.parametrize("channels", [1, 2, 3, 4])
@pytest.mark.parametrize("sixteen", [False, True])
def test_png_files_read_back(channels, sixteen):
    draw = np.random.default_rng(channels + 10 * sixteen)
    picture = draw.random((23, 31, channels)) * 255.0
    picture[0, 0] = 255.0
    data = native.png_bytes(picture if channels > 1 else picture[..., 0], sixteen=sixteen, dither="none",
                            compression=channels * 2)
    scale, top = (257, 65535) if sixteen else (1, 255)
    expected = np.array([exact_round(float(x), scale, top) for x in picture.reshape(-1)]).reshape(picture.shape)
    # The library's own reader.
    from test_png import read

    status, levels, bits, message, _ = read(data)
    assert status == 0 and message == "" and bits == (16 if sixteen else 8)
    assert np.array_equal(levels.astype(np.int64), expected)
    # Pillow's, for the modes it keeps exactly.
    if not sixteen:
        shown = np.asarray(Image.open(io.BytesIO(data))).astype(np.int64).reshape(expected.shape)
        assert np.array_equal(shown, expected)
    elif channels == 1:
        shown = np.asarray(Image.open(io.BytesIO(data))).astype(np.int64).reshape(expected.shape)
        assert np.array_equal(shown, expected)


def test_png_icc_profiles_go_where_they_go():
    picture = np.full((4, 6, 3), 100.0)
    good = profile()
    data = native.png_bytes(picture, icc_profile=good)
    assert Image.open(io.BytesIO(data)).info.get("icc_profile") == good
    # A grey profile with an RGB picture is left out, and said so.
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        data = native.png_bytes(picture, icc_profile=profile(b"GRAY"))
    assert any("ICC profile is not written" in str(w.message) for w in caught)
    assert Image.open(io.BytesIO(data)).info.get("icc_profile") is None
    # A profile that compresses into fewer than 92 bytes of chunk is stored uncompressed.
    small = profile(size=132)
    small = small[:128] + struct.pack(">I", 0)
    data = native.png_bytes(np.full((2, 2, 3), 5.0), icc_profile=small)
    at = data.index(b"iCCP")
    length = struct.unpack(">I", data[at - 4:at])[0]
    assert length >= 92
    assert Image.open(io.BytesIO(data)).info.get("icc_profile") == small


def test_png_refusals():
    with pytest.raises(ValueError):
        native.png_bytes(np.zeros((2, 2)), compression=10)
    with pytest.raises(native.UnroundError) as caught:
        native.png_bytes(np.zeros((2, 2, 5)))
    assert caught.value.status == native.Status.OPTIONS
    with pytest.raises(native.UnroundError):
        native.png_bytes(np.array([[math.nan]]))


tiff_tags = helpers.tiff_tags


@pytest.mark.parametrize("channels,ycbcr", [(1, False), (2, False), (3, False), (3, True), (4, False)])
def test_tiff_holds_the_samples_bit_for_bit(channels, ycbcr):
    draw = np.random.default_rng(channels)
    picture = draw.standard_normal((5, 7, channels)) * 1e3
    picture[0, 0, 0] = -0.0
    picture[1, 1, 0] = 5e-324
    given = picture if channels > 1 else picture[..., 0]
    data = native.tiff_bytes(given, ycbcr=ycbcr, icc_profile=profile(b"RGB " if channels > 2 else b"GRAY"))
    tags, _ = tiff_tags(data)
    assert tags[256][2] == 7 and tags[257][2] == 5 and tags[277][3][:2] == struct.pack("<H", channels)
    assert tags[259][3][:2] == struct.pack("<H", 1) and tags[284][3][:2] == struct.pack("<H", 1)
    photometric = struct.unpack("<H", tags[262][3][:2])[0]
    assert photometric == (6 if ycbcr else (2 if channels >= 3 else 1))
    offset, length = tags[273][2], tags[279][2]
    assert length == 5 * 7 * channels * 8
    samples = np.frombuffer(data[offset:offset + length], dtype="<f8").reshape(5, 7, channels)
    assert samples.tobytes() == picture.astype("<f8").tobytes()
    assert (338 in tags) == (channels in (2, 4))
    assert (532 in tags) == ycbcr
    kind, n, at, _ = tags[34675]
    assert data[at:at + n] == profile(b"RGB " if channels > 2 else b"GRAY")
    if channels > 2:
        bits = struct.unpack(f"<{channels}H", data[tags[258][2]:tags[258][2] + 2 * channels])
        formats = struct.unpack(f"<{channels}H", data[tags[339][2]:tags[339][2] + 2 * channels])
    else:
        bits = struct.unpack(f"<{channels}H", tags[258][3][:2 * channels])
        formats = struct.unpack(f"<{channels}H", tags[339][3][:2 * channels])
    assert set(bits) == {64} and set(formats) == {3}
    with pytest.raises(native.UnroundError):
        native.tiff_bytes(np.zeros((2, 2, 2)), ycbcr=True)


def test_the_icc_check_names_what_it_refuses():
    library = native._lib()
    library.unround_jpeg_check_icc.restype = ctypes.c_int32
    library.unround_jpeg_check_icc.argtypes = (ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_char_p,
                                               ctypes.c_size_t)

    def check(made: bytes, channels: int = 3) -> str:
        reason = ctypes.create_string_buffer(200)
        taken = library.unround_jpeg_check_icc(made, len(made), channels, reason, 200)
        return "" if taken else reason.value.decode()

    good = profile()
    assert check(good) == "" and check(profile(b"GRAY"), 1) == ""
    assert "132" in check(good[:100])
    assert "header says" in check(good[:-4])
    assert "acsp" in check(good[:36] + b"xxxx" + good[40:])
    assert "color space" in check(good, 1)
    assert "class" in check(good[:12] + b"link" + good[16:])
    assert "connection" in check(good[:20] + b"RGB " + good[24:])
    assert "tags" in check(good[:128] + struct.pack(">I", 100) + good[132:])
    assert "lies outside" in check(good[:132] + b"desc" + struct.pack(">II", 190, 40) + good[144:])


@pytest.mark.parametrize("orientation", range(9))
def test_the_eight_orientations_are_numpy_s_turns(orientation):
    picture = np.arange(5 * 7 * 2, dtype=np.float64).reshape(5, 7, 2)
    expected = {
        0: picture, 1: picture, 2: picture[:, ::-1], 3: picture[::-1, ::-1], 4: picture[::-1],
        5: picture.transpose(1, 0, 2), 6: np.rot90(picture, k=-1), 7: np.rot90(picture, k=-1)[::-1],
        8: np.rot90(picture, k=1),
    }[orientation]
    assert np.array_equal(native.oriented(picture, orientation), expected)
    assert np.array_equal(native.oriented(picture[..., 0], orientation), expected[..., 0])
    with pytest.raises(native.UnroundError):
        native.oriented(picture, 9)
