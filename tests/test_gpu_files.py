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


def block_dct(plane: np.ndarray) -> np.ndarray:
    """The orthonormal 8 x 8 DCT of a plane's blocks, (rows, columns, 8, 8)."""
    k = np.arange(8)
    basis = np.sqrt(2 / 8) * np.cos(np.pi * (2 * k[None, :] + 1) * k[:, None] / 16)
    basis[0] /= np.sqrt(2)
    rows, columns = plane.shape[0] // 8, plane.shape[1] // 8
    blocks = plane.reshape(rows, 8, columns, 8).transpose(0, 2, 1, 3)
    return basis @ blocks @ basis.T


def test_a_picture_made_elsewhere_is_measured_and_projected():
    # docs/math.md 12: the canvas of Pillow's own decoding, extended by its edges; each
    # coefficient's excess in steps against the intervals; the projection's coefficients
    # clipped to them; and the projection of the projection moves nothing.
    made = picture(37, 45, 1, seed=5)
    data = jpeg(made, quality=60)
    shown = np.asarray(Image.open(io.BytesIO(data))).astype(np.float64)
    settings = ["--method", "mmse"]
    print("projection of Pillow's decoding of a 37x45 grey JPEG file", flush=True)
    projected = native.project(shown, data, settings)
    assert projected.result is None and projected.excess is not None
    problem = projected.problems[0]
    rows, columns = problem.lower.shape[:2]
    canvas = np.pad(shown, ((0, 8 * rows - 37), (0, 8 * columns - 45)), mode="edge") - 0.0
    coefficients = block_dct(canvas)
    steps = problem.steps.reshape(1, 1, 8, 8)
    expected = (np.maximum(problem.lower - coefficients, 0) + np.maximum(coefficients - problem.upper, 0)) / steps
    assert np.max(np.abs(projected.excess[0] - expected)) <= 1e-9
    assert np.count_nonzero(expected > 1e-9) > 0
    clipped = np.clip(coefficients, problem.lower, problem.upper)
    assert np.max(np.abs(projected.coefficients[0] - clipped)) <= 1e-9
    assert np.all(projected.coefficients[0] >= problem.lower) and np.all(projected.coefficients[0] <= problem.upper)
    # The projection of the projection's picture: within the intervals at the blocks that the
    # picture covers whole (docs/math.md 12: the samples beyond it moved too, and cut).
    again = native.project(projected.picture, data, settings)
    assert np.max(again.excess[0][: 37 // 8, : 45 // 8]) <= 1e-9


def test_a_colour_picture_is_projected_through_jfif_s_ycbcr():
    made = picture(32, 48, 3, seed=6)
    data = jpeg(made, quality=70, subsampling=2)
    shown = np.asarray(Image.open(io.BytesIO(data))).astype(np.float64)
    print("projection of a 32x48 4:2:0 JPEG file's decoding", flush=True)
    projected = native.project(shown, data, ["--method", "mmse"])
    assert projected.picture.shape == (32, 48, 3) and len(projected.excess) == 3
    for c, problem in enumerate(projected.problems):
        assert np.all(projected.coefficients[c] >= problem.lower) and np.all(projected.coefficients[c] <= problem.upper)
    again = native.project(projected.picture, data, ["--method", "mmse"])
    assert max(float(np.max(e)) for e in again.excess) <= 1e-6
