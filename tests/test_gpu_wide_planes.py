"""The 64-bit addressing of the kernels (WidePlanes), which only canvases too large for
buffer resources take, forced on small ones by UNROUND_WIDE_PLANES=1: the same values by
another route, so each solve must equal the buffer-resource solve bit for bit, for one
picture and for a batch, grey and 4:2:0, TV and TGV, binary32 and BF16 duals; and the
4:2:0 case stays within the oracle's reach as the buffer path does."""

import numpy as np
import pytest

import helpers
from unround import native, synthetic

pytestmark = pytest.mark.gpu


def grey(height, width, seed, quality):
    return synthetic.encode_grey(synthetic.grey_picture(height, width, seed), quality)


def colour(height, width, seed, quality):
    return synthetic.encode_colour420(synthetic.colour_picture(height, width, seed), quality)


def same(a, b, label):
    assert np.array_equal(a.canvas, b.canvas), f"{label}: canvas"
    for ours, theirs in zip(a.coefficients, b.coefficients):
        assert np.array_equal(ours, theirs), f"{label}: coefficients"
    assert np.array_equal(a.result.history.primal, b.result.history.primal), f"{label}: records"
    assert np.array_equal(a.result.history.dual, b.result.history.dual), f"{label}: records"


CASES = [
    ("tv grey", lambda: grey(72, 104, 21, 50), ["--method", "tv", "--iterations", "40", "--tolerance", "0"]),
    ("tgv grey", lambda: grey(96, 136, 22, 40), ["--method", "tgv", "--iterations", "40", "--tolerance", "0"]),
    ("tgv 4:2:0", lambda: colour(136, 200, 23, 75), ["--method", "tgv", "--iterations", "30", "--tolerance", "0"]),
    ("tgv 4:2:0 mixed", lambda: colour(80, 96, 24, 60),
     ["--method", "tgv", "--iterations", "30", "--tolerance", "0", "--precision", "mixed"]),
    ("tv 4:2:0 apart", lambda: colour(64, 88, 25, 50),
     ["--method", "tv", "--channels", "apart", "--iterations", "30", "--tolerance", "0"]),
]


@pytest.mark.parametrize("label,make,options", CASES, ids=[case[0] for case in CASES])
def test_wide_planes_equal_buffer_planes(monkeypatch, label, make, options):
    image = make()
    print(f"wide planes: {label} {image.height}x{image.width}", flush=True)
    monkeypatch.delenv("UNROUND_WIDE_PLANES", raising=False)
    narrow = native.solve(image, options)
    monkeypatch.setenv("UNROUND_WIDE_PLANES", "1")
    wide = native.solve(image, options)
    same(wide, narrow, label)


def test_wide_planes_batch_equals_buffer_planes(monkeypatch):
    images = [grey(96, 104, 31, 50), colour(120, 88, 32, 75), colour(136, 200, 33, 90)]
    options = ["--method", "tgv", "--iterations", "40", "--tolerance", "0", "--relative-tolerance", "2e-3"]
    print("wide planes: batch of 3", flush=True)
    monkeypatch.delenv("UNROUND_WIDE_PLANES", raising=False)
    narrow, _ = native.solve_batch(images, options)
    monkeypatch.setenv("UNROUND_WIDE_PLANES", "1")
    wide, _ = native.solve_batch(images, options)
    for index, (a, b) in enumerate(zip(wide, narrow)):
        assert a.result.iterations == b.result.iterations
        same(a, b, f"batch image {index}")


def test_wide_planes_against_the_oracle(monkeypatch):
    picture = synthetic.colour_picture(136, 200, 5)
    image = synthetic.encode_colour420(picture, 75)
    options = ["--method", "tgv", "--iterations", "60", "--tolerance", "0"]
    monkeypatch.setenv("UNROUND_WIDE_PLANES", "1")
    print("wide planes: 136x200 4:2:0 against the oracle", flush=True)
    decoded = native.solve(image, options)
    expected = helpers.oracle_solve(image, options)
    ours, theirs = decoded.planes, expected.canvas[:, :136, :200]
    a, b = helpers.rounded(ours), helpers.rounded(theirs)
    assert int(np.max(np.abs(a - b))) <= 1
    assert float(np.mean(a != b)) <= 1e-3
    relative = np.abs(decoded.result.history.primal - expected.primal) / np.abs(expected.primal)
    assert np.max(relative) < 1e-4
