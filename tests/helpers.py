"""What the tests share: the package and the oracle on the same inputs, and the measures of
agreement (8-bit rounding as the reference's output.rs:99-107, PSNR, requantization)."""

from __future__ import annotations

import numpy as np

from oracle import oracle


def oracle_components(image) -> list[oracle.Component]:
    return [
        oracle.Component(levels=np.asarray(c.coefficients, dtype=np.int16), table=np.asarray(c.quant_table).reshape(64),
                         h=int(c.h_samp_factor), v=int(c.v_samp_factor))
        for c in image.components
    ]


def oracle_solve(image, options: list[str], start: dict | None = None) -> oracle.OracleResult:
    return oracle.solve(int(image.height), int(image.width), int(image.color_space), oracle_components(image), options,
                        start)


def rounded(samples: np.ndarray) -> np.ndarray:
    """8-bit levels: the nearest, halves away from 0, clamped to 0-255."""
    values = np.sign(samples) * np.floor(np.abs(samples) + 0.5)
    return np.clip(values, 0, 255).astype(np.int64)


def psnr(picture: np.ndarray, original: np.ndarray) -> float:
    error = np.mean((np.clip(picture, 0, 255) - original.astype(np.float64)) ** 2)
    return float(10 * np.log10(255.0**2 / error))


def block_dct(plane: np.ndarray) -> np.ndarray:
    """The orthonormal 8x8 DCT of every block, (rows, columns, 8, 8), from the oracle's basis."""
    basis = oracle.basis()
    h, w = plane.shape
    blocks = plane.reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3)
    return np.einsum("km,abmn,un->abku", basis, blocks, basis)


def requantized(canvas: np.ndarray, image, index: int, cells: tuple[int, int], rows: int, columns: int) -> np.ndarray:
    """The levels of the first rows x columns blocks of a component of a canvas: its cells'
    means' DCT less the level shift, over the steps, rounded to the nearest integer."""
    component = image.components[index]
    down, across = cells
    plane = canvas[: rows * 8 * down, : columns * 8 * across]
    if down * across > 1:
        plane = plane.reshape(rows * 8, down, columns * 8, across).mean(axis=(1, 3))
    coefficients = block_dct(plane)
    coefficients[:, :, 0, 0] -= 1024.0
    table = np.asarray(component.quant_table, dtype=np.float64).reshape(8, 8)
    return np.rint(coefficients / table).astype(np.int64)


def kept_levels(image, index: int, rows: int, columns: int) -> np.ndarray:
    return np.asarray(image.components[index].coefficients, dtype=np.int64)[:rows, :columns]


def oracle_unrolled_inputs(result: oracle.OracleResult) -> tuple[np.ndarray, np.ndarray]:
    """The start state and problem of a grey oracle solve (sides multiples of 8) in the unrolled
    restatement's layout: start (8, H, W) = x (the inverse DCT of the centres clipped into
    their intervals), w = grad x, p = r = 0; problem (rows, columns, 64, 4) = lower, upper,
    centre and the prox weight of every coefficient (the flat blocks' weights where marked)."""
    import torch

    import unrolled_reference

    lower, upper, centres = result.lower[0], result.upper[0], result.centres[0]
    rows, columns = lower.shape[:2]
    flat = result.flat[0].astype(bool)[:, :, None, None]
    weights = np.where(flat, result.flat_weights[0][None, None], result.weights[0][None, None])
    problem = np.stack([lower, upper, centres, np.broadcast_to(weights, lower.shape)], axis=-1)
    zeta = torch.as_tensor(np.clip(centres, lower, upper)[None], dtype=torch.float64)
    x = unrolled_reference._idct(zeta, unrolled_reference.basis())
    start = np.zeros((8, 8 * rows, 8 * columns))
    start[0] = x[0].numpy()
    start[1] = unrolled_reference._forward_x(x)[0].numpy()
    start[2] = unrolled_reference._forward_y(x)[0].numpy()
    return start, problem.reshape(rows, columns, 64, 4)


def parity_check(image, original, options, precision="single", ratio_limit=1e-3, psnr_limit=0.01):
    """The GPU's solve of `image` against the oracle's under the same options (see
    parity_compare)."""
    from unround import native

    print(f"parity {image.height}x{image.width} space {int(image.color_space)} {' '.join(options)} {precision}",
          flush=True)
    expected = oracle_solve(image, options)
    decoded = native.solve(image, options + ["--precision", precision])
    return parity_compare(decoded, expected, image, original, options, ratio_limit, psnr_limit)


# Output samples the binary64 result puts within this distance of a rounding boundary
# (n + 1/2) are decided by rounding noise, not by the method: a block with only its DC
# coefficient (level l, step q) has samples 128 + l q / 8, an exact half wherever
# l q = 4 (mod 8), and binary64's own inverse DCT lands on either side of it. They are
# held to the 1-level bound only.
TIE_BAND = 1e-3


def parity_compare(decoded, expected, image, original, options, ratio_limit=1e-3, psnr_limit=0.01):
    """A GPU result against the oracle's of the same image and options, on the output pixels
    (RGB for colour, as 8-bit output is written; the plane for grey): 8-bit values within 1
    level, differing on at most `ratio_limit` of the samples (those within TIE_BAND of a
    rounding boundary aside), PSNR against `original` (if given) within `psnr_limit` dB,
    every coefficient inside its interval and (without a slack) requantizing to its level,
    the records' iterations equal and their primal values within 1e-4. Returns (decoded,
    expected)."""
    import unround

    C, H, W = expected.canvas.shape
    assert decoded.canvas.shape == (C, H, W)
    h, w = int(image.height), int(image.width)
    ours_planes = decoded.planes
    theirs_planes = expected.canvas[:, :h, :w]
    worst = float(np.max(np.abs(ours_planes - theirs_planes)))
    if C == 1:
        ours, theirs = ours_planes[0], theirs_planes[0]
    else:
        ours, theirs = unround.to_rgb(ours_planes), unround.to_rgb(theirs_planes)
    a, b = rounded(ours), rounded(theirs)
    tied = np.abs(np.abs(theirs - np.floor(theirs)) - 0.5) <= TIE_BAND
    differing = float(np.mean((a != b) & ~tied))
    largest = int(np.max(np.abs(a - b)))
    print(f"  max |diff| {worst:.3g} (planes), 8-bit differing {differing:.5f} "
          f"(and {float(np.mean((a != b) & tied)):.5f} at ties, {float(np.mean(tied)):.5f} tied), largest {largest}")
    assert largest <= 1
    assert differing <= ratio_limit
    if original is not None:
        ours_psnr, theirs_psnr = psnr(ours, original), psnr(theirs, original)
        print(f"  PSNR {ours_psnr:.4f} vs {theirs_psnr:.4f}")
        assert abs(ours_psnr - theirs_psnr) < psnr_limit
    widened = "--slack" in options
    for c, problem in enumerate(decoded.problems):
        rows, columns = decoded.coefficients[c].shape[:2]
        assert np.all(decoded.coefficients[c] > problem.lower) and np.all(decoded.coefficients[c] < problem.upper)
        if widened:
            continue  # a slack widens the intervals beyond the file's own: no requantization
        levels = requantized(decoded.canvas[c], image, c, problem.ratio, rows, columns)
        assert np.array_equal(levels, kept_levels(image, c, rows, columns)), f"component {c} requantizes"
    history = decoded.result.history
    assert list(history.iterations) == list(expected.record_iterations)
    relative = np.abs(history.primal - expected.primal) / np.abs(expected.primal)
    print(f"  records: primal rel {np.max(relative):.3g}")
    assert np.max(relative) < 1e-4
    return decoded, expected


def tiff_tags(data: bytes) -> tuple[dict, int]:
    """The one IFD of a little-endian TIFF file: tag -> (type, count, value or offset, the
    entry's 4 value bytes), checked ascending and alone; and its offset."""
    import struct

    assert data[:4] == b"II*\x00"
    ifd = struct.unpack("<I", data[4:8])[0]
    count = struct.unpack("<H", data[ifd:ifd + 2])[0]
    tags = {}
    previous = -1
    for k in range(count):
        tag, kind, n, value = struct.unpack("<HHII", data[ifd + 2 + 12 * k:ifd + 14 + 12 * k])
        assert tag > previous
        previous = tag
        tags[tag] = (kind, n, value, data[ifd + 10 + 12 * k:ifd + 14 + 12 * k])
    assert struct.unpack("<I", data[ifd + 2 + 12 * count:ifd + 6 + 12 * count])[0] == 0
    return tags, ifd
