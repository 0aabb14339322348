"""The oracle's pictures of rounded samples (docs/math.md, 11) against the reference's own
checks of them (rust/unround/tests/samples.rs): the intervals exactly, at every level of 1,
2, 4, 8 and 16 bits against the rational ends they stand for; the proximal map's values
and optimality; the conjugate by Fenchel and Young; the solvers on a smooth picture rounded
to 8 bits; and the defaults of when they stop."""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle


def smooth(channels: int, height: int, width: int) -> tuple[np.ndarray, np.ndarray]:
    """The smooth picture of rust/unround/tests/samples.rs:22-45 and its nearest 8-bit
    levels, halves away from 0 (all its values are positive)."""
    c, y, x = np.meshgrid(np.arange(channels, dtype=np.float64), np.arange(height, dtype=np.float64),
                          np.arange(width, dtype=np.float64), indexing="ij")
    values = 40.0 + 30.0 * c + 0.23 * x + 0.11 * y + 0.002 * (x - 20.0) * (x - 20.0)
    whole = np.floor(values)
    levels = np.clip(whole + (values - whole >= 0.5), 0, 255).astype(np.uint16)
    return values, levels


def row(levels):
    return np.asarray(levels, dtype=np.uint16).reshape(1, 1, -1)


def down(value: float) -> float:
    return math.nextafter(value, -math.inf)


def up(value: float) -> float:
    return math.nextafter(value, math.inf)


def test_intervals_are_half_a_level_on_each_side():
    # rust/unround/tests/samples.rs:48-83
    low, high, centre = oracle.samples_intervals([0, 1, 128, 254, 255], 8)
    assert low.tolist() == [0.0, 0.5, 127.5, 253.5, 254.5]
    assert high.tolist() == [down(0.5), down(1.5), down(128.5), down(254.5), 255.0]
    assert centre.tolist() == [0.0, 1.0, 128.0, 254.0, 255.0]
    low, high, centre = oracle.samples_intervals([0, 3, 15], 4)
    assert low.tolist() == [0.0, 42.5, 246.5]
    assert high.tolist() == [down(8.5), down(59.5), 255.0]
    assert centre.tolist() == [0.0, 51.0, 255.0]
    low, high, centre = oracle.samples_intervals([1000], 16)
    assert Fraction(low[0]) > Fraction(1999 * 255, 131070)
    assert Fraction(high[0]) < Fraction(2001 * 255, 131070)
    assert centre[0] == 1000.0 * 255.0 / 65535.0
    low, high, _ = oracle.samples_intervals([0, 100, 255], 8, slack=0.25)
    assert low.tolist() == [0.0, 99.25, 254.25]
    assert high.tolist() == [0.75, 100.75, 255.0]
    with pytest.raises(ValueError):
        oracle.samples_intervals([256], 8)
    with pytest.raises(ValueError):
        oracle.samples_intervals([1], 3)
    with pytest.raises(ValueError):
        oracle.samples_intervals([1], 8, slack=-1.0)


def rounded(value: float, scale: int, top: int) -> int:
    """The level a sample rounds to at `scale` times it, halves away from 0, in 0..top,
    exactly."""
    whole = math.floor(Fraction(value) * scale + Fraction(1, 2))
    return min(max(whole, 0), top)


@pytest.mark.parametrize("bits", [1, 2, 4, 8, 16])
def test_every_interval_holds_the_doubles_that_round_to_its_level(bits):
    # rust/unround/tests/samples.rs:93-156: the least double at or above (v - 1/2) s and the
    # greatest below (v + 1/2) s, with s = 255 / (2^bits - 1); 0 and 255 at the ends.
    top = (1 << bits) - 1
    levels = np.arange(top + 1)
    low, high, _ = oracle.samples_intervals(levels, bits)
    for level in range(top + 1):
        if level == 0:
            assert low[0] == 0.0
        else:
            end = Fraction((2 * level - 1) * 255, 2 * top)
            assert Fraction(low[level]) >= end and Fraction(down(low[level])) < end, (bits, level)
        if level == top:
            assert high[level] == 255.0
        else:
            end = Fraction((2 * level + 1) * 255, 2 * top)
            assert Fraction(high[level]) < end and Fraction(up(high[level])) >= end, (bits, level)
    if bits not in (8, 16):
        return
    # Rounded to the file's own bits (8: the sample itself, 16: 257 times it), the ends give
    # back their level and the doubles just beyond them the levels beside it.
    scale = 1 if bits == 8 else 257
    for level in range(top + 1):
        assert rounded(low[level], scale, top) == level
        assert rounded(high[level], scale, top) == level
        if level > 0:
            assert rounded(down(low[level]), scale, top) == level - 1
        if level < top:
            assert rounded(up(high[level]), scale, top) == level + 1


def test_the_proximal_map_is_optimal_and_the_projection_idempotent():
    # rust/unround/tests/samples.rs:158-195
    levels = [10, 20, 30, 40]
    values = np.array([9.0, 20.25, 31.0, -4.0])
    out = oracle.samples_prox(levels, 8, values, 0.7)
    assert out.tolist() == [9.5, 20.25, down(30.5), 39.5]
    assert oracle.samples_prox(levels, 8, out, 0.7).tolist() == out.tolist()
    weighted = oracle.samples_prox(levels, 8, values, 0.5, mu=2.0)
    low, high, centre = oracle.samples_intervals(levels, 8)
    for z, v, c, a, b in zip(weighted, values, centre, low, high):
        gradient = (z - v) + (z - c)
        if a < z < b:
            assert gradient == 0.0
        elif z == a:
            assert gradient >= 0.0
        else:
            assert gradient <= 0.0


@pytest.mark.parametrize("mu", [0.0, 0.5, 3.0])
def test_the_conjugate_meets_fenchel_and_young(mu):
    # rust/unround/tests/samples.rs:197-232
    levels = row([0, 7, 200, 255])
    low, high, _ = oracle.samples_intervals(levels, 8)
    for s in (-3.0, -0.25, 0.0, 0.125, 2.0):
        conjugate = oracle.samples_conjugate(levels, 8, np.full((1, 1, 4), s), mu=mu)
        best = -math.inf
        for step in range(65):
            c = low + step / 64.0 * (high - low)
            best = max(best, float(np.sum(s * c)) - oracle.samples_value(levels, 8, c.reshape(1, 1, 4), mu=mu))
        assert conjugate >= best - 1e-9 * (1 + abs(best)), (s, conjugate, best)
        if mu == 0.0:
            assert abs(conjugate - best) <= 1e-9 * (1 + abs(best)), (s, conjugate, best)


def test_the_value_is_half_mu_times_the_squared_distance_to_the_centres():
    # G(x) = mu / 2 * sum (x - c)^2 at samples within the intervals, 0 where mu is
    # (rust/unround/src/samples.rs:202-220).
    levels = row([10, 20, 30])
    _, _, centre = oracle.samples_intervals(levels, 8)
    x = centre + np.array([0.25, -0.5, 0.0]).reshape(1, 1, 3)
    assert oracle.samples_value(levels, 8, x, mu=2.0) == 0.25**2 + 0.5**2
    assert oracle.samples_value(levels, 8, x, mu=0.0) == 0.0


def rms(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


@pytest.mark.parametrize("method", ["tv", "tgv"])
def test_the_solvers_find_the_smooth_picture_within_the_intervals(method):
    # rust/unround/tests/samples.rs:240-272 (TGV's error 0.07 against the levels' 0.29 there)
    values, levels = smooth(3, 24, 40)
    low, high, centre = oracle.samples_intervals(levels, 8)
    result = oracle.solve_samples(levels, 8, ["--method", method, "--iterations", "4000"])
    picture = result.picture
    assert np.all(picture >= low) and np.all(picture <= high)
    assert oracle.samples_prox(levels, 8, picture, 1.0).tolist() == picture.tolist()
    for primal, dual in zip(result.primal, result.dual):
        assert dual <= primal * (1 + 1e-12) + 1e-9
    assert result.iterations > 0
    if method == "tgv":
        error, levels_error = rms(values, picture), rms(values, centre)
        print(f"tgv rms {error:.4f} against the levels' {levels_error:.4f}")
        assert error < 0.1 and error < levels_error / 2


@pytest.mark.parametrize("method,weight,ratio", [("tv", "--alpha", 10.0), ("tgv", "--alpha1", 0.3)])
def test_the_solvers_take_the_defaults_of_rounded_samples(method, weight, ratio):
    # rust/unround/tests/samples.rs:274-314: tolerance 1e-4 times the first-order weight,
    # 10000 iterations and the ratio over the weight squared; without weight scaling the
    # plain tolerance and ratio.
    _, levels = smooth(1, 12, 16)
    base = ["--method", method, weight, "2"]
    found = oracle.solve_samples(levels, 8, base)
    given = oracle.solve_samples(levels, 8, base + ["--tolerance", "2e-4", "--iterations", "10000",
                                                    "--step-ratio", repr(ratio / 4.0)])
    assert found.stop == 1  # converged
    assert found.iterations == given.iterations
    assert np.array_equal(found.picture, given.picture)
    found = oracle.solve_samples(levels, 8, base + ["--no-weight-scaling"])
    given = oracle.solve_samples(levels, 8, base + ["--no-weight-scaling", "--tolerance", "1e-4", "--iterations",
                                                    "10000", "--step-ratio", repr(ratio)])
    assert found.iterations == given.iterations
    assert np.array_equal(found.picture, given.picture)
