"""Pictures of rounded samples on the GPU (unround_solve_samples, the solve behind a PNG
file's reconstruction) against the oracle on the same levels: in binary64 the picture and
every record to rounding; in binary32 within the tolerance of the iterate's type and, on
the reference's own smooth picture (rust/unround/tests/samples.rs:240-272), TGV's error
below 0.1 and half the levels'. Every result lies within the exact intervals and rounds
back to its levels; the alpha is solved alone; the defaults of rounded samples stop where
the oracle's do."""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle
from unround import native

pytestmark = pytest.mark.gpu


def smooth(channels: int, height: int, width: int) -> tuple[np.ndarray, np.ndarray]:
    """The reference's smooth picture (rust/unround/tests/samples.rs:22-45), (C, H, W), and its
    8-bit levels."""
    c, y, x = np.meshgrid(np.arange(channels, dtype=np.float64), np.arange(height, dtype=np.float64),
                          np.arange(width, dtype=np.float64), indexing="ij")
    values = 40.0 + 30.0 * c + 0.23 * x + 0.11 * y + 0.002 * (x - 20.0) * (x - 20.0)
    whole = np.floor(values)
    return values, np.clip(whole + (values - whole >= 0.5), 0, 255).astype(np.uint16)


def noisy(channels: int, height: int, width: int, bits: int, seed: int) -> np.ndarray:
    """Levels of a smooth field with noise, (C, H, W), of `bits` bits."""
    draw = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    top = (1 << bits) - 1
    planes = []
    for c in range(channels):
        field = 0.5 + 0.35 * np.sin(x / (7.0 + c)) * np.cos(y / 9.0) + 0.05 * draw.standard_normal((height, width))
        planes.append(np.clip(np.rint(field * top), 0, top))
    return np.stack(planes).astype(np.uint16)


def interleaved(planes: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(planes, 0, -1))


def gpu(planes: np.ndarray, bits: int, options: list[str], observer=None) -> native.Decoded:
    given = interleaved(planes)
    return native.solve_samples(given[..., 0] if given.shape[-1] == 1 else given, bits, options, observer=observer)


def planes_of(decoded: native.Decoded) -> np.ndarray:
    return decoded.planes


def rounded(values: np.ndarray, bits: int) -> np.ndarray:
    """The levels the samples round to at the file's bits (8: the sample, 16: 257 times it),
    halves away from 0, computed exactly (rust/unround/src/output.rs:93-107)."""
    scale = 1 if bits == 8 else 257
    top = (1 << bits) - 1
    flat = [min(max(math.floor(Fraction(float(v)) * scale + Fraction(1, 2)), 0), top) for v in np.ravel(values)]
    return np.asarray(flat, dtype=np.int64).reshape(values.shape)


def assert_within(picture: np.ndarray, levels: np.ndarray, bits: int, slack: float = 0.0) -> None:
    low, high, _ = oracle.samples_intervals(levels, bits, slack)
    assert np.all(picture >= low) and np.all(picture <= high)
    assert oracle.samples_prox(levels, bits, picture, 1.0, slack=slack).tolist() == picture.tolist()


def records_match(decoded: native.Decoded, expected: oracle.SamplesResult, relative: float) -> None:
    """The records of the iterations both took, within `relative` of the primal value."""
    history = decoded.result.history
    count = min(len(history.iterations), len(expected.record_iterations))
    assert count > 1
    assert history.iterations[:count].tolist() == expected.record_iterations[:count].tolist()
    scale = np.maximum(np.abs(expected.primal[:count]), 1.0)
    assert np.max(np.abs(history.primal[:count] - expected.primal[:count]) / scale) <= relative
    assert np.max(np.abs(history.dual[:count] - expected.dual[:count]) / scale) <= relative


@pytest.mark.parametrize("method", ["tv", "tgv"])
def test_binary64_matches_the_oracle(method):
    _, levels = smooth(3, 24, 40)
    options = ["--method", method, "--iterations", "400", "--tolerance", "0", "--precision", "double"]
    print(f"samples binary64 {method} 3x24x40, 400 iterations", flush=True)
    decoded = gpu(levels, 8, options)
    expected = oracle.solve_samples(levels, 8, options)
    difference = float(np.max(np.abs(planes_of(decoded) - expected.picture)))
    print(f"max |gpu - oracle| {difference:.3g}", flush=True)
    assert difference <= 1e-8
    records_match(decoded, expected, 1e-10)
    assert len(decoded.result.history.iterations) == len(expected.record_iterations)
    assert decoded.input_format == "png" and decoded.bits == 8 and decoded.coefficients == ()
    assert decoded.result.iterations == expected.iterations == 400
    assert_within(planes_of(decoded), levels, 8)


def test_binary32_finds_the_smooth_picture():
    # rust/unround/tests/samples.rs:240-272: TGV's error 0.07 against the levels' 0.29.
    values, levels = smooth(3, 24, 40)
    options = ["--method", "tgv", "--iterations", "4000"]
    print("samples binary32 tgv 3x24x40, 4000 iterations", flush=True)
    decoded = gpu(levels, 8, options)
    expected = oracle.solve_samples(levels, 8, options)
    picture = planes_of(decoded)
    assert_within(picture, levels, 8)
    error = float(np.sqrt(np.mean((picture - values) ** 2)))
    levels_error = float(np.sqrt(np.mean((levels - values) ** 2)))
    print(f"rms to the picture {error:.4f} (levels {levels_error:.4f}); max |gpu - oracle| "
          f"{np.max(np.abs(picture - expected.picture)):.3g}", flush=True)
    assert error < 0.1 and error < levels_error / 2
    assert np.max(np.abs(picture - expected.picture)) <= 2e-2
    # binary32 iterates: the feasible dual's xi is a difference of differences of r~, whose
    # rounding moves the dual value by a few 1e-4 of the primal's.
    records_match(decoded, expected, 1e-3)
    for primal, dual in zip(decoded.result.history.primal, decoded.result.history.dual):
        assert dual <= primal * (1 + 1e-6) + 1e-6


@pytest.mark.parametrize("bits", [1, 2, 4, 8, 16])
def test_every_depth_rounds_back_to_its_levels(bits):
    levels = noisy(1, 40, 70, bits, seed=bits)
    options = ["--method", "tgv", "--iterations", "200", "--tolerance", "0"]
    print(f"samples {bits} bits tgv 40x70, 200 iterations", flush=True)
    decoded = gpu(levels, bits, options)
    expected = oracle.solve_samples(levels, bits, options)
    picture = planes_of(decoded)
    assert_within(picture, levels, bits)
    if bits in (8, 16):
        assert np.array_equal(rounded(picture, bits), levels.astype(np.int64))
    # Each sample moves within an interval of 255 / (2^bits - 1) wide; binary32 tracks the oracle
    # to a small part of it.
    width = 255.0 / ((1 << bits) - 1)
    assert np.max(np.abs(picture - expected.picture)) <= max(2e-2, 1e-3 * width)
    assert decoded.bits == bits


def test_weight_and_slack_match_the_oracle():
    levels = noisy(3, 33, 47, 8, seed=5)
    options = ["--method", "tgv", "--mu", "0.5", "--slack", "0.25", "--iterations", "300", "--tolerance", "0",
               "--precision", "double",
               "--channels", "apart", "--channel-weights", "1,0.5,0.75", "--partial-radius", "20"]
    print("samples mu 0.5 slack 0.25 apart, weights, partial gap, binary64", flush=True)
    decoded = gpu(levels, 8, options)
    expected = oracle.solve_samples(levels, 8, options)
    assert float(np.max(np.abs(planes_of(decoded) - expected.picture))) <= 1e-8
    records_match(decoded, expected, 1e-10)
    partial = decoded.result.history.partial_gap
    assert np.allclose(partial, expected.partial, rtol=1e-9, atol=1e-9)
    assert_within(planes_of(decoded), levels, 8, slack=0.25)


def test_the_alpha_is_solved_alone_and_the_defaults_stop_as_the_oracle_does():
    colours = noisy(3, 24, 32, 8, seed=11)
    alpha = smooth(1, 24, 32)[1]
    levels = np.concatenate([colours, alpha])
    options = ["--method", "tgv", "--precision", "double", "--channel-weights", "1,0.5,0.5"]
    seen: list[tuple[int, int]] = []
    decoded = gpu(levels, 8, options, observer=lambda record: seen.append((record.canvas.shape[0], record.iteration)))
    expected_colours = oracle.solve_samples(colours, 8, options)
    expected_alpha = oracle.solve_samples(alpha, 8, ["--method", "tgv", "--precision", "double"])
    print(f"colours {decoded.result.iterations} (oracle {expected_colours.iterations}), alpha "
          f"{decoded.alpha_result.iterations} (oracle {expected_alpha.iterations})", flush=True)
    assert int(decoded.result.stop) == expected_colours.stop and int(decoded.alpha_result.stop) == expected_alpha.stop
    assert decoded.alpha_result.stop == native.Stop.CONVERGED
    assert decoded.result.iterations == expected_colours.iterations
    assert decoded.alpha_result.iterations == expected_alpha.iterations
    assert np.max(np.abs(decoded.planes[:3] - expected_colours.picture)) <= 1e-8
    assert np.max(np.abs(decoded.planes[3:] - expected_alpha.picture)) <= 1e-8
    assert decoded.picture.shape == (24, 32, 4)
    assert np.array_equal(decoded.picture, np.moveaxis(decoded.planes, 0, -1))
    assert np.array_equal(decoded.canvas, decoded.planes)
    # The observer sees the colours' records, then the alpha's.
    colour_records = len(decoded.result.history.iterations) - 1
    alpha_records = len(decoded.alpha_result.history.iterations) - 1
    assert [channels for channels, _ in seen] == [3] * colour_records + [1] * alpha_records


def test_an_observer_stops_the_solve():
    _, levels = smooth(1, 16, 24)
    decoded = gpu(levels, 8, ["--method", "tv"], observer=lambda record: record.iteration >= 30)
    assert decoded.result.stop == native.Stop.OBSERVER and decoded.result.iterations == 30
    assert_within(planes_of(decoded), levels, 8)


def test_a_png_file_decodes_as_its_levels_solve():
    # unround_decode of a PNG file (rust/capi/src/lib.rs:550-554): the file layer's levels,
    # then the solve of solve_samples, bit for bit; the ICC profile kept; a tRNS chunk warned of.
    import io
    import struct
    import warnings

    from PIL import Image

    colours = noisy(3, 20, 28, 8, seed=21)
    alpha = smooth(1, 20, 28)[1]
    rgba = interleaved(np.concatenate([colours, alpha])).astype(np.uint8)
    profile = bytearray(160)
    profile[0:4] = struct.pack(">I", 160)
    profile[36:40] = b"acsp"
    buffer = io.BytesIO()
    Image.fromarray(rgba, mode="RGBA").save(buffer, "PNG", icc_profile=bytes(profile))
    options = ["--method", "tgv", "--iterations", "120", "--tolerance", "0"]
    print("decode of a 20x28 RGBA PNG file, 120 iterations", flush=True)
    decoded = native.decode(buffer.getvalue(), options)
    direct = native.solve_samples(rgba, 8, options)
    assert decoded.input_format == "png" and decoded.bits == 8
    assert np.array_equal(decoded.picture, direct.picture) and np.array_equal(decoded.planes, direct.planes)
    assert decoded.icc_profile == bytes(profile)
    assert decoded.alpha_result is not None and decoded.alpha_result.iterations == 120
    assert int(decoded.color_space) == 3
    # An RGB file with a tRNS chunk: reconstructed, and the colour not taken is said.
    rgb = np.ascontiguousarray(rgba[..., :3])
    buffer = io.BytesIO()
    Image.fromarray(rgb, mode="RGB").save(buffer, "PNG", transparency=(1, 2, 3))
    data = buffer.getvalue()
    assert b"tRNS" in data
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        decoded = native.decode(data, options)
    assert any("tRNS" in str(w.message) for w in caught)
    assert decoded.alpha_result is None and decoded.picture.shape == (20, 28, 3)
    # What a PNG file does not take is refused before it is read.
    with pytest.raises(native.UnroundError) as refused:
        native.decode(data, ["--method", "mmse"])
    assert refused.value.status == native.Status.OPTIONS
