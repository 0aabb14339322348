"""Seeded synthetic pictures, quantized as a baseline JPEG encoder does: the workloads of the
benchmarks and the inputs of the parity tests.

A greyscale picture of seed s is, in binary64 and then rounded to 8 bits:
  * 128 plus six cosine fields, each of amplitude 40 / 3 times U[0.5, 1], frequencies
    U[0, 3) cycles per picture down and across and a phase U[0, 2 pi);
  * twelve axis-aligned rectangles or ellipses (even odds), their centres uniform on the
    picture, their half-sides U[0.03, 0.2] of the picture's sides, each painted over what
    is below with a constant intensity U[0, 255];
  * Gaussian noise of sigma 3; clipped to [0, 255].
A colour picture takes one such luminance base L (seed s) and a field of its own for each
channel (seeds 1000 s + 1, 2, 3): R, G, B = clip(0.75 L + 0.25 F_c), rounded.

Quantization (docs/math.md:124-127): the level of a coefficient is the nearest integer,
ties to even, to the block DCT of the samples less 128 over the step; the steps are the
Annex K tables of ITU-T T.81 scaled by the IJG rule (quality q below 50: 5000 / q, else
200 - 2q percent; each step clamp((base * scale + 50) // 100, 1, 255)). Colour is JFIF's
YCbCr, each channel rounded to 8 bits; 4:2:0 chroma are the means of 2 x 2 cells
rounded to integers; the picture is filled out to whole MCUs by repeating its last row
and column, as libjpeg does.
"""

from __future__ import annotations

import numpy as np

from unround import jpegio

__all__ = [
    "CHROMA",
    "LUMA",
    "colour_picture",
    "encode_colour420",
    "encode_grey",
    "grey_picture",
    "ijg_table",
]

# ITU-T T.81 Annex K, Tables K.1 and K.2, natural order.
LUMA = np.array(
    [16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56, 14, 17, 22, 29, 51,
     87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101,
     72, 92, 95, 98, 112, 100, 103, 99],
    dtype=np.int64,
)
CHROMA = np.array(
    [17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99,
     99, 99, 99] + [99] * 32,
    dtype=np.int64,
)


def ijg_table(base: np.ndarray, quality: int) -> np.ndarray:
    """The IJG scaling of a base table to a quality from 1 to 100."""
    quality = min(max(int(quality), 1), 100)
    scale = 5000 // quality if quality < 50 else 200 - 2 * quality
    return np.clip((base * scale + 50) // 100, 1, 255).astype(np.uint16)


def _dct_matrix() -> np.ndarray:
    k = np.arange(8)[:, None]
    n = np.arange(8)[None, :]
    c = np.cos(np.pi * (2 * n + 1) * k / 16) * 0.5
    c[0, :] = np.sqrt(1.0 / 8.0)
    return c


_C = _dct_matrix()


def grey_picture(height: int, width: int, seed: int) -> np.ndarray:
    """The seeded greyscale picture, uint8 (height, width)."""
    return np.clip(np.rint(_field(height, width, seed)), 0, 255).astype(np.uint8)


def _field(height: int, width: int, seed: int, noise: bool = True) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rows = np.arange(height)[:, None] / height
    columns = np.arange(width)[None, :] / width
    picture = np.full((height, width), 128.0)
    for _ in range(6):
        amplitude = 40.0 / 3.0 * rng.uniform(0.5, 1.0)
        fy, fx = rng.uniform(0.0, 3.0, size=2)
        phase = rng.uniform(0.0, 2 * np.pi)
        picture += amplitude * np.cos(2 * np.pi * (fy * rows + fx * columns) + phase)
    for _ in range(12):
        cy, cx = rng.uniform(0.0, 1.0, size=2)
        hy, hx = rng.uniform(0.03, 0.2, size=2)
        intensity = rng.uniform(0.0, 255.0)
        if rng.uniform() < 0.5:
            mask = (np.abs(rows - cy) <= hy) & (np.abs(columns - cx) <= hx)
        else:
            mask = ((rows - cy) / hy) ** 2 + ((columns - cx) / hx) ** 2 <= 1.0
        picture = np.where(mask, intensity, picture)
    if noise:
        picture = picture + rng.normal(0.0, 3.0, size=(height, width))
    return np.clip(picture, 0.0, 255.0)


def colour_picture(height: int, width: int, seed: int) -> np.ndarray:
    """The seeded RGB picture, uint8 (height, width, 3), its channels sharing a luminance base."""
    base = _field(height, width, seed)
    channels = [0.75 * base + 0.25 * _field(height, width, 1000 * seed + c + 1) for c in range(3)]
    return np.clip(np.rint(np.stack(channels, axis=-1)), 0, 255).astype(np.uint8)


def _padded(plane: np.ndarray, rows: int, columns: int) -> np.ndarray:
    return np.pad(plane, ((0, rows - plane.shape[0]), (0, columns - plane.shape[1])), mode="edge")


def _levels(plane: np.ndarray, table: np.ndarray) -> np.ndarray:
    """Levels (rows, columns, 8, 8) of a plane whose sides are multiples of 8."""
    h, w = plane.shape
    blocks = plane.astype(np.float64).reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3) - 128.0
    coefficients = np.einsum("km,abmn,un->abku", _C, blocks, _C)
    return np.rint(coefficients / table.reshape(8, 8).astype(np.float64)).astype(np.int16)


def requantized(canvas: np.ndarray, image: jpegio.Image, index: int, cells: tuple[int, int] = (1, 1)) -> np.ndarray:
    """The levels that a reconstruction's canvas plane (H, W) of component `index` quantizes
    to: its cells' means, block DCT less the level shift, over the steps, rounded to the
    nearest integer; (rows, columns, 8, 8) like the component's own levels."""
    component = image.components[index]
    rows, columns = component.coefficients.shape[:2]
    down, across = cells
    plane = np.asarray(canvas, dtype=np.float64)[: rows * 8 * down, : columns * 8 * across]
    if down * across > 1:
        plane = plane.reshape(rows * 8, down, columns * 8, across).mean(axis=(1, 3))
    return _levels(plane, np.asarray(component.quant_table))


def encode_grey(picture: np.ndarray, quality: int) -> jpegio.Image:
    """A greyscale picture quantized at a quality with the scaled luminance table."""
    height, width = picture.shape
    table = ijg_table(LUMA, quality)
    plane = _padded(picture, -(-height // 8) * 8, -(-width // 8) * 8)
    component = jpegio.make_component(_levels(plane, table), table, 1, 1)
    return jpegio.make_image(height, width, jpegio.ColorSpace.GRAYSCALE, (component,))


def _ycbcr_planes(picture: np.ndarray) -> list[np.ndarray]:
    """JFIF's Y, Cb and Cr of an RGB picture, rounded to 8 bits as an encoder stores them."""
    rgb = picture.astype(np.float64)
    y = 0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]
    cb = 128.0 + (rgb[..., 2] - y) / 1.772
    cr = 128.0 + (rgb[..., 0] - y) / 1.402
    return [np.clip(np.rint(p), 0, 255) for p in (y, cb, cr)]


def encode_colour444(picture: np.ndarray, quality: int) -> jpegio.Image:
    """An RGB picture as JFIF YCbCr without subsampling, quantized at a quality with the
    scaled luma table for Y and chroma table for Cb and Cr."""
    height, width, _ = picture.shape
    rows, columns = -(-height // 8) * 8, -(-width // 8) * 8
    tables = (ijg_table(LUMA, quality), ijg_table(CHROMA, quality), ijg_table(CHROMA, quality))
    components = tuple(jpegio.make_component(_levels(_padded(plane, rows, columns), table), table, 1, 1)
                       for plane, table in zip(_ycbcr_planes(picture), tables))
    return jpegio.make_image(height, width, jpegio.ColorSpace.YCBCR, components)


def encode_colour420(picture: np.ndarray, quality: int) -> jpegio.Image:
    """An RGB picture as JFIF YCbCr 4:2:0 quantized at a quality with the scaled tables."""
    height, width, _ = picture.shape
    planes = _ycbcr_planes(picture)
    rows, columns = -(-height // 16) * 16, -(-width // 16) * 16
    luma = _padded(planes[0], rows, columns)
    chroma = []
    for plane in planes[1:]:
        full = _padded(plane, rows, columns)
        means = full.reshape(rows // 2, 2, columns // 2, 2).mean(axis=(1, 3))
        chroma.append(np.floor(means + 0.5))
    luma_table, chroma_table = ijg_table(LUMA, quality), ijg_table(CHROMA, quality)
    components = (
        jpegio.make_component(_levels(luma, luma_table), luma_table, 2, 2),
        jpegio.make_component(_levels(chroma[0], chroma_table), chroma_table, 1, 1),
        jpegio.make_component(_levels(chroma[1], chroma_table), chroma_table, 1, 1),
    )
    return jpegio.make_image(height, width, jpegio.ColorSpace.YCBCR, components)


def batch_pictures(count: int = 64) -> tuple[list[jpegio.Image], list[np.ndarray], list[str]]:
    """The mixed batch: picture k (seeds 1..count) of height and width uniform on the multiples
    of 8 in [96, 1920], quality uniform in [30, 90], grey or 4:2:0 colour with even odds, each
    drawn from its own generator (seed 1000 + k); with the uncompressed pictures and the kinds
    ("grey" or "colour")."""
    images, originals, kinds = [], [], []
    for seed in range(1, count + 1):
        draw = np.random.default_rng(1000 + seed)
        height, width = (int(v) * 8 for v in draw.integers(12, 241, size=2))
        quality = int(draw.integers(30, 91))
        if draw.random() < 0.5:
            picture = grey_picture(height, width, seed)
            images.append(encode_grey(picture, quality))
            kinds.append("grey")
        else:
            picture = colour_picture(height, width, seed)
            images.append(encode_colour420(picture, quality))
            kinds.append("colour")
        originals.append(picture)
    return images, originals, kinds
