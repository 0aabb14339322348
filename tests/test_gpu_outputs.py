"""JPEG file to output file through the GPU: unround_decode of the committed JPEG files
(TGV, 60 iterations, binary64 on the GPU), then the library's writers.

- The 8-bit PNG (each sample rounded) equals, byte for byte, the expected file that
  tests/golden/make_outputs.py wrote from the oracle's decode of the same file through
  libpng 1.6, the reference's PNG writer (c/jpegio/src/png.c:250-276). The pictures have
  no sample within 1e-6 of a rounding boundary, so a GPU decode within binary64's reach of
  the oracle's rounds to the same levels.
- The TIFF holds the decoded picture's binary64 samples bit for bit, with the tags the
  reference's writer sets (rust/unround/tests/tiff.rs:134-161), and those samples are
  within 1e-8 of the oracle's."""

from __future__ import annotations

import struct
from pathlib import Path

import numpy as np
import pytest

import helpers
from unround import jpegio, native

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "outputs"
SETTINGS = ["--method", "tgv", "--iterations", "60", "--tolerance", "0", "--precision", "double"]


@pytest.mark.parametrize("name", ["grey_q50", "colour420_q75"])
def test_decode_to_png_and_tiff(name):
    data = (GOLDEN / f"{name}.jpg").read_bytes()
    print(f"decode to files: {name}", flush=True)
    decoded = native.decode(data, SETTINGS)
    picture = decoded.picture
    height, width = picture.shape[:2]
    channels = 1 if picture.ndim == 2 else picture.shape[2]

    assert native.png_bytes(picture, dither="none") == (GOLDEN / f"{name}_tgv60.png").read_bytes()

    image = jpegio.read(data)
    expected = helpers.oracle_solve(image, SETTINGS[:-2])
    planes = expected.canvas[:, :height, :width]
    theirs = planes[0] if channels == 1 else native.to_rgb(planes)
    assert float(np.max(np.abs(picture - theirs))) <= 1e-8

    tiff = native.tiff_bytes(picture)
    tags, _ = helpers.tiff_tags(tiff)
    assert (tags[256][2], tags[257][2], tags[278][2]) == (width, height, height)
    assert tags[277][3][:2] == struct.pack("<H", channels)
    assert tags[259][3][:2] == struct.pack("<H", 1) and tags[284][3][:2] == struct.pack("<H", 1)
    assert struct.unpack("<H", tags[262][3][:2])[0] == (2 if channels == 3 else 1)
    assert 338 not in tags and 34675 not in tags
    offset, length = tags[273][2], tags[279][2]
    assert length == height * width * channels * 8
    samples = np.frombuffer(tiff[offset:offset + length], dtype="<f8")
    assert samples.tobytes() == np.ascontiguousarray(picture, dtype="<f8").tobytes()
