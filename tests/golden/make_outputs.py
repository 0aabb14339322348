"""Writes the expected files of the GPU decode-to-file test (tests/test_gpu_outputs.py):
two small JPEG files (PIL's encoder: grey q50, and RGB q75 with 4:2:0 chroma), and the 8-bit
PNG files of their TGV decodes (60 iterations, no tolerance, binary64) as the oracle solves
them: its canvas cut to the picture, JFIF's RGB for colour, each sample rounded (halves
away from 0, clamped to 0..255, rust/unround/src/output.rs:99-107) and written by libpng 1.6
as the reference writes PNG (c/jpegio/src/png.c:250-276). A picture with a sample within
1e-6 of a rounding boundary is refused (its 8-bit level would be decided by rounding noise);
the seed moves on until none is.

Run from the repository root: python tests/golden/make_outputs.py"""

from __future__ import annotations

import io
import sys
from pathlib import Path

import numpy as np
from PIL import Image

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path[:0] = [str(ROOT), str(ROOT / "jpeg-unround_amd"), str(ROOT / "tests")]

import helpers  # noqa: E402
from libpng import LIBPNG, libpng_file  # noqa: E402
from unround import jpegio, native  # noqa: E402

SETTINGS = ["--method", "tgv", "--iterations", "60", "--tolerance", "0"]
CASES = {"grey_q50": (44, 52, 1, {"quality": 50}), "colour420_q75": (48, 56, 3, {"quality": 75, "subsampling": 2})}


def picture(height: int, width: int, channels: int, seed: int) -> np.ndarray:
    draw = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    planes = [128 + 70 * np.sin((x + 5 * c) / 11.0) * np.cos(y / 7.0) + draw.normal(0, 4, (height, width))
              for c in range(channels)]
    made = np.clip(np.stack(planes, axis=-1), 0, 255).astype(np.uint8)
    return made[..., 0] if channels == 1 else made


def oracle_picture(data: bytes) -> np.ndarray:
    image = jpegio.read(data)
    result = helpers.oracle_solve(image, SETTINGS)
    planes = result.canvas[:, :image.height, :image.width]
    return planes[0] if planes.shape[0] == 1 else native.to_rgb(planes)


def main() -> None:
    if LIBPNG is None:
        raise SystemExit("libpng 1.6 is not installed here")
    out = HERE / "outputs"
    out.mkdir(exist_ok=True)
    for name, (height, width, channels, options) in CASES.items():
        for seed in range(1, 100):
            buffer = io.BytesIO()
            Image.fromarray(picture(height, width, channels, seed)).save(buffer, "JPEG", **options)
            data = buffer.getvalue()
            expected = oracle_picture(data)
            distance = np.abs(np.abs(expected - np.floor(expected)) - 0.5)
            if float(np.min(distance)) > 1e-6:
                break
        else:
            raise SystemExit(f"{name}: every seed has a sample at a rounding boundary")
        (out / f"{name}.jpg").write_bytes(data)
        (out / f"{name}_tgv60.png").write_bytes(libpng_file(helpers.rounded(expected), 8, -1))
        print(f"{name}: seed {seed}, {len(data)} bytes of JPEG, nearest boundary {float(np.min(distance)):.3g}")


if __name__ == "__main__":
    main()
