"""The JPEG reader of the file layer (unround_jpegio_read, unround_jpegio_decode_planes;
c/jpegio/include/unround/jpegio.h:188-213) without a GPU:

* files written here from known coefficients (baseline Huffman, interleaved and not, every
  sampling, restart intervals) read back to those coefficients exactly;
* files that Pillow (libjpeg-turbo) writes, baseline and progressive with optimized
  tables, read to the same coefficients whichever the coding (the quantized DCT does not
  depend on it), and their planes by the accurate integer inverse DCT equal to Pillow's
  own decoding of the luma (and of every plane of 4:4:4 files) bit for bit;
* the metadata: EXIF orientation, ICC profile over several APP2 markers, the colour space
  libjpeg infers; and what it refuses or warns of: a cut file, limits, 12-bit and
  arithmetic frames.
"""

from __future__ import annotations

import io
import struct

import numpy as np
import pytest
from PIL import Image

from unround import jpegio, native

NATURAL = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7,
           14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46,
           53, 60, 61, 54, 47, 55, 62, 63]


# ---------------------------------------------------------------------------------------
# A baseline encoder of given coefficients (T.81 F.1.2), with simple, valid tables: the 12
# DC sizes at 4 bits, every AC symbol at 8.


class Bits:
    def __init__(self) -> None:
        self.out = bytearray()
        self.value = 0
        self.count = 0

    def put(self, value: int, length: int) -> None:
        for shift in range(length - 1, -1, -1):
            self.value = (self.value << 1) | ((value >> shift) & 1)
            self.count += 1
            if self.count == 8:
                self.out.append(self.value)
                if self.value == 0xFF:
                    self.out.append(0)
                self.value, self.count = 0, 0

    def flush(self) -> None:
        while self.count:
            self.put(1, 1)


def size_of(value: int) -> int:
    return abs(value).bit_length()


def bits_of(value: int, size: int) -> int:
    return value if value >= 0 else value + (1 << size) - 1


AC_SYMBOLS = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]
DC_CODES = {s: (s, 4) for s in range(12)}
AC_CODES = {symbol: (k, 8) for k, symbol in enumerate(AC_SYMBOLS)}


def dht(kind: int, slot: int, symbols: list[int], length: int) -> bytes:
    counts = [0] * 16
    counts[length - 1] = len(symbols)
    body = bytes([(kind << 4) | slot]) + bytes(counts) + bytes(symbols)
    return b"\xff\xc4" + struct.pack(">H", 2 + len(body)) + body


def encode_block(bits: Bits, block: np.ndarray, previous: int) -> int:
    zigzag = [int(block.reshape(64)[NATURAL[k]]) for k in range(64)]
    difference = zigzag[0] - previous
    size = size_of(difference)
    bits.put(*DC_CODES[size])
    bits.put(bits_of(difference, size), size)
    run = 0
    for k in range(1, 64):
        value = zigzag[k]
        if value == 0:
            run += 1
            continue
        while run > 15:
            bits.put(*AC_CODES[0xF0])
            run -= 16
        size = size_of(value)
        bits.put(*AC_CODES[(run << 4) | size])
        bits.put(bits_of(value, size), size)
        run = 0
    if run:
        bits.put(*AC_CODES[0x00])
    return zigzag[0]


def jpeg_file(height: int, width: int, components: list[tuple[int, int, np.ndarray, np.ndarray]],
              restart: int = 0, interleaved: bool = True, extra: bytes = b"") -> bytes:
    """A baseline file of components (h, v, coefficients (rows, columns, 8, 8) padded to whole
    MCUs, table (64,) natural order)."""
    out = bytearray(b"\xff\xd8") + extra
    for slot, (_, _, _, table) in enumerate(components):
        body = bytes([slot]) + bytes(int(table[NATURAL[k]]) for k in range(64))
        out += b"\xff\xdb" + struct.pack(">H", 2 + len(body)) + body
    frame = struct.pack(">BHHB", 8, height, width, len(components))
    for index, (h, v, _, _) in enumerate(components):
        frame += bytes([index + 1, (h << 4) | v, index])
    out += b"\xff\xc0" + struct.pack(">H", 2 + len(frame)) + frame
    out += dht(0, 0, list(range(12)), 4) + dht(1, 0, AC_SYMBOLS, 8)
    if restart:
        out += b"\xff\xdd" + struct.pack(">HH", 4, restart)
    max_h = max(c[0] for c in components)
    max_v = max(c[1] for c in components)
    scans = [list(range(len(components)))] if interleaved else [[c] for c in range(len(components))]
    for scan in scans:
        header = bytes([len(scan)]) + b"".join(bytes([c + 1, 0x00]) for c in scan) + b"\x00\x3f\x00"
        out += b"\xff\xda" + struct.pack(">H", 2 + len(header)) + header
        bits = Bits()
        predictors = [0] * len(components)
        units: list[list[tuple[int, int, int]]] = []
        if len(scan) == 1:
            h, v, levels, _ = components[scan[0]]
            across = -(-width * h // (8 * max_h))
            down = -(-height * v // (8 * max_v))
            units = [[(scan[0], row, column)] for row in range(down) for column in range(across)]
        else:
            for mcu_row in range(-(-height // (8 * max_v))):
                for mcu_column in range(-(-width // (8 * max_h))):
                    units.append([(c, mcu_row * components[c][1] + b, mcu_column * components[c][0] + a)
                                  for c in scan for b in range(components[c][1]) for a in range(components[c][0])])
        for number, unit in enumerate(units):
            if restart and number > 0 and number % restart == 0:
                bits.flush()
                out += bits.out
                out += bytes([0xFF, 0xD0 + ((number // restart - 1) % 8)])
                bits = Bits()
                predictors = [0] * len(components)
            for c, row, column in unit:
                predictors[c] = encode_block(bits, components[c][2][row, column], predictors[c])
        bits.flush()
        out += bits.out
    return bytes(out + b"\xff\xd9")


def random_levels(rows: int, columns: int, seed: int) -> np.ndarray:
    draw = np.random.default_rng(seed)
    levels = np.zeros((rows, columns, 8, 8), dtype=np.int16)
    levels[..., 0, 0] = draw.integers(-60, 60, size=(rows, columns))
    mask = draw.random((rows, columns, 8, 8)) < 0.25
    levels += (mask * draw.integers(-30, 31, size=(rows, columns, 8, 8))).astype(np.int16)
    levels[0, 0, 7, 7] = 1023
    levels[-1, -1, 0, 1] = -1023
    return levels


def table(seed: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(1, 100, size=64).astype(np.uint16)


@pytest.mark.parametrize("sampling,size", [((1, 1), (24, 40)), ((2, 2), (37, 29)), ((2, 1), (16, 51)),
                                           ((1, 2), (33, 20))])
@pytest.mark.parametrize("restart", [0, 1, 3])
@pytest.mark.parametrize("interleaved", [True, False])
def test_written_coefficients_read_back_exactly(sampling, size, restart, interleaved):
    h, v = sampling
    height, width = size
    mcus_down, mcus_across = -(-height // (8 * v)), -(-width // (8 * h))
    luma = random_levels(mcus_down * v, mcus_across * h, seed=restart + 10 * h + 100 * v)
    chroma = [random_levels(mcus_down, mcus_across, seed=s) for s in (7, 8)]
    components = [(h, v, luma, table(1)), (1, 1, chroma[0], table(2)), (1, 1, chroma[1], table(3))]
    data = jpeg_file(height, width, components, restart=restart, interleaved=interleaved)
    image = jpegio.read(data)
    assert image.warning == "" and image.warnings == 0 and not image.progressive
    assert (image.width, image.height, image.color_space) == (width, height, jpegio.ColorSpace.YCBCR)
    assert (image.max_h_samp_factor, image.max_v_samp_factor) == (h, v)
    for c, (ch, cv, levels, quant) in enumerate(components):
        got = image.components[c]
        rows, columns = -(-height * cv // (8 * v)), -(-width * ch // (8 * h))
        assert got.coefficients.shape == (rows, columns, 8, 8)
        assert np.array_equal(got.coefficients, levels[:rows, :columns])
        assert np.array_equal(got.quant_table.reshape(64), quant)
        assert (got.h_samp_factor, got.v_samp_factor, got.id) == (ch, cv, c + 1)


def dct_levels(plane: np.ndarray, quant: np.ndarray) -> np.ndarray:
    """The quantized orthonormal block DCT of a plane (rows and columns multiples of 8)."""
    k = np.arange(8)
    basis = np.sqrt(2 / 8) * np.cos(np.pi * (2 * k[None, :] + 1) * k[:, None] / 16)
    basis[0] /= np.sqrt(2)
    rows, columns = plane.shape[0] // 8, plane.shape[1] // 8
    blocks = (plane.astype(np.float64) - 128).reshape(rows, 8, columns, 8).transpose(0, 2, 1, 3)
    coefficients = basis @ blocks @ basis.T
    return np.rint(coefficients / quant.reshape(8, 8)).astype(np.int16)


@pytest.mark.parametrize("restart", [0, 2, 5])
def test_written_files_decode_as_pillow_decodes_them(restart):
    # Realistic coefficients (a smooth picture's), with restart intervals: Pillow's decoding
    # of the luma equals the planes of the layer's accurate integer inverse DCT.
    picture = smooth_picture(48, 64, 3, seed=restart)
    quant = np.full(64, 4, dtype=np.uint16)
    quant[:8] = 2
    components = [(1, 1, dct_levels(picture[..., c], quant), quant) for c in range(3)]
    data = jpeg_file(48, 64, components, restart=restart)
    image = jpegio.read(data)
    for c in range(3):
        assert np.array_equal(image.components[c].coefficients, components[c][2])
    shown = Image.open(io.BytesIO(data))
    shown.draft("YCbCr", shown.size)
    ycbcr = np.asarray(shown)
    planes = jpegio.decode_planes(data).planes
    for c in range(3):
        assert np.array_equal(planes[c], ycbcr[..., c])


def pillow_file(picture: np.ndarray, **options) -> bytes:
    buffer = io.BytesIO()
    Image.fromarray(picture).save(buffer, "JPEG", **options)
    return buffer.getvalue()


def smooth_picture(height: int, width: int, channels: int, seed: int) -> np.ndarray:
    draw = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    base = 128 + 60 * np.sin(x / 9.0) * np.cos(y / 13.0)
    planes = [base + 25 * np.cos((x + 3 * c) / (5.0 + c)) + draw.normal(0, 6, (height, width)) for c in range(channels)]
    picture = np.clip(np.stack(planes, axis=-1), 0, 255).astype(np.uint8)
    return picture[..., 0] if channels == 1 else picture


@pytest.mark.parametrize("quality", [20, 75, 95])
def test_pillow_grey_files_baseline_and_progressive(quality):
    picture = smooth_picture(61, 83, 1, seed=quality)
    baseline = pillow_file(picture, quality=quality)
    progressive = pillow_file(picture, quality=quality, progressive=True, optimize=True)
    one, other = jpegio.read(baseline), jpegio.read(progressive)
    assert other.progressive and not one.progressive
    assert one.color_space == jpegio.ColorSpace.GRAYSCALE
    assert np.array_equal(one.components[0].coefficients, other.components[0].coefficients)
    assert np.array_equal(one.components[0].quant_table, other.components[0].quant_table)
    for data in (baseline, progressive):
        shown = np.asarray(Image.open(io.BytesIO(data)))
        assert np.array_equal(jpegio.decode_planes(data).planes[0][:61, :83], shown)


@pytest.mark.parametrize("subsampling", [0, 1, 2])
def test_pillow_colour_files_baseline_and_progressive(subsampling):
    picture = smooth_picture(45, 70, 3, seed=subsampling)
    baseline = pillow_file(picture, quality=80, subsampling=subsampling)
    progressive = pillow_file(picture, quality=80, subsampling=subsampling, progressive=True, optimize=True)
    one, other = jpegio.read(baseline), jpegio.read(progressive)
    assert one.color_space == jpegio.ColorSpace.YCBCR
    expected_h = {0: (1, 1), 1: (2, 1), 2: (2, 2)}[subsampling]
    assert (one.components[0].h_samp_factor, one.components[0].v_samp_factor) == expected_h
    for a, b in zip(one.components, other.components):
        assert np.array_equal(a.coefficients, b.coefficients)
    shown = Image.open(io.BytesIO(baseline))
    shown.draft("YCbCr", shown.size)
    planes = jpegio.decode_planes(baseline).planes
    ycbcr = np.asarray(shown)
    assert np.array_equal(planes[0][:45, :70], ycbcr[..., 0])
    if subsampling == 0:
        for c in (1, 2):
            assert np.array_equal(planes[c][:45, :70], ycbcr[..., c])


def test_metadata_orientation_icc_and_colour_space():
    picture = smooth_picture(20, 30, 3, seed=4)
    exif = b"Exif\x00\x00MM\x00*\x00\x00\x00\x08\x00\x01\x01\x12\x00\x03\x00\x00\x00\x01\x00\x06\x00\x00\x00\x00\x00\x00"
    profile = bytes(range(256)) * 300  # 76800 bytes: two APP2 markers
    data = pillow_file(picture, exif=exif, icc_profile=profile)
    assert data.count(b"ICC_PROFILE\x00") == 2
    image = jpegio.read(data)
    assert image.exif_orientation == 6 and image.icc_profile == profile
    little = b"Exif\x00\x00II*\x00\x08\x00\x00\x00\x01\x00\x12\x01\x03\x00\x01\x00\x00\x00\x03\x00\x00\x00\x00\x00\x00\x00"
    assert jpegio.read(pillow_file(picture, exif=little)).exif_orientation == 3
    # RGB kept as RGB: an Adobe marker of transform 0.
    rgb = jpegio.read(pillow_file(picture, keep_rgb=True))
    assert rgb.color_space == jpegio.ColorSpace.RGB
    # Without a JFIF marker, component identifiers 'R', 'G', 'B' mean RGB.
    levels = [random_levels(3, 4, seed=s) for s in range(3)]
    file_ = jpeg_file(24, 32, [(1, 1, levels[c], table(c)) for c in range(3)])
    assert jpegio.read(file_).color_space == jpegio.ColorSpace.YCBCR
    renamed = file_.replace(b"\x01\x11\x00\x02\x11\x01\x03\x11\x02", b"R\x11\x00G\x11\x01B\x11\x02")
    renamed = renamed.replace(b"\x03\x01\x00\x02\x00\x03\x00\x00\x3f\x00", b"\x03R\x00G\x00B\x00\x00\x3f\x00")
    assert jpegio.read(renamed).color_space == jpegio.ColorSpace.RGB


def test_cut_and_refused_files():
    picture = smooth_picture(40, 48, 1, seed=9)
    data = pillow_file(picture, quality=60)
    # Cut within the scan: the rest is taken as zeros, with libjpeg's warning.
    end = data.index(b"\xff\xda") + (len(data) - data.index(b"\xff\xda")) // 2
    cut = jpegio.read(data[:end])
    assert cut.warnings >= 1 and cut.warning == "Premature end of JPEG file"
    whole = jpegio.read(data).components[0].coefficients
    assert np.array_equal(cut.components[0].coefficients[:2], whole[:2])
    with pytest.raises(jpegio.JpegioError) as caught:
        jpegio.read(data[:end], warnings_are_errors=True)
    assert caught.value.kind == jpegio.ErrorKind.DECODE
    with pytest.raises(jpegio.JpegioError) as caught:
        jpegio.read(data, max_pixels=40 * 48 - 1)
    assert caught.value.kind == jpegio.ErrorKind.LIMIT
    assert jpegio.read(data, max_pixels=40 * 48).width == 48
    progressive = pillow_file(picture, quality=60, progressive=True)
    scans = progressive.count(b"\xff\xda")
    with pytest.raises(jpegio.JpegioError) as caught:
        jpegio.read(progressive, max_scans=scans - 1)
    assert caught.value.kind == jpegio.ErrorKind.LIMIT and "scans" in caught.value.message
    assert jpegio.read(progressive, max_scans=scans).progressive
    at = data.index(b"\xff\xc0")
    for code, kind in ((0xC9, jpegio.ErrorKind.UNSUPPORTED), (0xC3, jpegio.ErrorKind.UNSUPPORTED)):
        with pytest.raises(jpegio.JpegioError) as caught:
            jpegio.read(data[:at + 1] + bytes([code]) + data[at + 2:])
        assert caught.value.kind == kind
    twelve = bytearray(data)
    twelve[at + 4] = 12
    with pytest.raises(jpegio.JpegioError) as caught:
        jpegio.read(bytes(twelve))
    assert caught.value.kind == jpegio.ErrorKind.UNSUPPORTED and "12-bit" in caught.value.message
    with pytest.raises(jpegio.JpegioError) as caught:
        jpegio.read(b"GIF89a")
    assert caught.value.kind == jpegio.ErrorKind.DECODE
    # Extraneous bytes before a marker: a warning, and the coefficients still read.
    at = data.index(b"\xff\xd9")
    extra = jpegio.read(data[:at] + b"\x12\x34" + data[at:])
    assert extra.warning.startswith("Corrupt JPEG data: 2 extraneous bytes before marker 0xd9")
    assert np.array_equal(extra.components[0].coefficients, jpegio.read(data).components[0].coefficients)


def test_decode_reads_a_jpeg_file_and_refuses_what_it_cannot_read():
    # unround_decode of a JPEG file goes through the reader; a file it cannot read is
    # UNROUND_ERROR_READ (rust/unround/src/error.rs:46-50).
    with pytest.raises(native.UnroundError) as caught:
        native.decode(b"\xff\xd8\xff", ["--method", "tgv"])
    assert caught.value.status == native.Status.READ
    with pytest.raises(native.UnroundError) as caught:
        native.decode(b"not a picture", ["--method", "tgv"])
    assert caught.value.status == native.Status.READ


def test_projection_refusals_come_before_the_gpu():
    # rust/capi/src/lib.rs:708-752: a PNG file, a picture of another size or other channels.
    picture = smooth_picture(16, 24, 1, seed=2)
    data = pillow_file(picture, quality=60)
    png = io.BytesIO()
    Image.fromarray(picture).save(png, "PNG")
    cases = [(png.getvalue(), np.zeros((16, 24)), native.Status.OPTIONS, "PNG"),
             (data, np.zeros((16, 23)), native.Status.OPTIONS, "16 x 24"),
             (data, np.zeros((16, 24, 3)), native.Status.OPTIONS, "channels"),
             (data[:20], np.zeros((16, 24)), native.Status.READ, "")]
    for source, given, status, words in cases:
        with pytest.raises(native.UnroundError) as caught:
            native.project(given, source, ["--method", "mmse"])
        assert caught.value.status == status and words in caught.value.message, caught.value.message
