"""Files through the GPU solve: unround_decode of JPEG files (the layer's reader, then the
solve of their components) against the solve of the same components read separately and
against the oracle; and the result's metadata (EXIF orientation, ICC profile, warnings)."""

from __future__ import annotations

import io

import numpy as np
import pytest
from PIL import Image

from oracle import oracle
from unround import jpegio, native

pytestmark = pytest.mark.gpu


def picture(height: int, width: int, channels: int, seed: int) -> np.ndarray:
    draw = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    planes = [128 + 70 * np.sin((x + 5 * c) / 11.0) * np.cos(y / 7.0) + draw.normal(0, 4, (height, width))
              for c in range(channels)]
    made = np.clip(np.stack(planes, axis=-1), 0, 255).astype(np.uint8)
    return made[..., 0] if channels == 1 else made


def jpeg(made: np.ndarray, **options) -> bytes:
    buffer = io.BytesIO()
    Image.fromarray(made).save(buffer, "JPEG", **options)
    return buffer.getvalue()


@pytest.mark.parametrize("channels,options", [(1, {"quality": 50}), (3, {"quality": 75, "subsampling": 2}),
                                              (3, {"quality": 90, "subsampling": 0, "progressive": True})])
def test_decode_of_a_jpeg_file_is_the_solve_of_its_components(channels, options):
    exif = b"Exif\x00\x00II*\x00\x08\x00\x00\x00\x01\x00\x12\x01\x03\x00\x01\x00\x00\x00\x08\x00\x00\x00\x00\x00\x00\x00"
    data = jpeg(picture(37, 45, channels, seed=channels), exif=exif, **options)
    settings = ["--method", "tv", "--iterations", "40", "--tolerance", "0", "--precision", "double"]
    print(f"decode of a {channels}-channel JPEG file {options}", flush=True)
    decoded = native.decode(data, settings)
    image = jpegio.read(data)
    direct = native.solve(image, settings)
    assert np.array_equal(decoded.canvas, direct.canvas) and np.array_equal(decoded.picture, direct.picture)
    assert decoded.exif_orientation == 8 and decoded.input_format == "jpeg" and decoded.bits == 8
    components = [oracle.Component(np.asarray(c.coefficients), np.asarray(c.quant_table).reshape(64),
                                   c.h_samp_factor, c.v_samp_factor) for c in image.components]
    expected = oracle.solve(image.height, image.width, int(image.color_space), components, settings)
    assert np.max(np.abs(decoded.canvas - expected.canvas)) <= 1e-8
