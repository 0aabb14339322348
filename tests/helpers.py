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
