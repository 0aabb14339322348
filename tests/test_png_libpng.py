"""The library's PNG writer against libpng 1.6, which writes the reference's PNG files
(c/jpegio/src/png.c:250-276: every setting libpng's own but the zlib level): the system's
libpng16 (with the same zlib the library links), driven through ctypes with the settings
png.c gives it, must write the same bytes for the same 8-bit or 16-bit levels. Skipped
where no libpng16 is installed."""

from __future__ import annotations

import numpy as np
import pytest

from libpng import LIBPNG, libpng_file
from unround import native

pytestmark = pytest.mark.skipif(LIBPNG is None, reason="no libpng 1.6 on this machine")


def levels_of(shape, seed, top):
    draw = np.random.default_rng(seed)
    y, x = np.mgrid[0:shape[0], 0:shape[1]]
    base = 0.5 + 0.35 * np.sin(x / 9.0) * np.cos(y / 5.0)
    planes = base if len(shape) == 2 else np.stack([base * (0.6 + 0.2 * k) for k in range(shape[2])], axis=-1)
    return np.clip(np.round(planes * top + draw.normal(0, top / 80, planes.shape)), 0, top)


SHAPES = [(1, 1), (2, 3, 3), (5, 7, 4), (37, 45, 3), (64, 64), (16, 2000, 2), (200, 300, 3), (3, 5000), (600, 700, 3)]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_eight_bit_files_are_libpng_s(shape):
    levels = levels_of(shape, sum(shape), 255)
    assert native.png_bytes(levels, dither="none") == libpng_file(levels, 8, -1)


@pytest.mark.parametrize("shape", SHAPES[:7], ids=["x".join(map(str, s)) for s in SHAPES[:7]])
def test_sixteen_bit_files_are_libpng_s(shape):
    levels = levels_of(shape, sum(shape) + 1, 65535)
    assert native.png_bytes(levels / 257.0, sixteen=True, dither="none") == libpng_file(levels, 16, -1)


@pytest.mark.parametrize("compression", [0, 1, 6, 9])
def test_every_level_is_libpng_s(compression):
    levels = levels_of((90, 130, 3), compression, 255)
    assert native.png_bytes(levels, dither="none", compression=compression) == libpng_file(levels, 8, compression)
