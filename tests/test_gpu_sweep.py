"""The fused sweep (the dual step of one iteration and the proximal step of the next in one
launch, taken with UNROUND_SWEEP=1) against the two kernels, the default: the same
arithmetic at every sample, so the canvases, coefficients, records, stopping iterations and
reasons are equal bit for bit. The sweep makes two samples of p once more where another work-group
writes them (the row above each run, the column left of each work-group); they must come
out as that work-group's own, so the result must not depend on how the rows are cut into
runs either (UNROUND_DUAL_ROWS: 16 against 64), as the reference's results do not depend
on its threads' strips (cpp/tests/test_pdhg.cpp:328-386)."""

import numpy as np
import pytest

from unround import native, synthetic

pytestmark = pytest.mark.gpu


def grey(height, width, seed, quality):
    return synthetic.encode_grey(synthetic.grey_picture(height, width, seed), quality)


def colour(height, width, seed, quality):
    return synthetic.encode_colour420(synthetic.colour_picture(height, width, seed), quality)


def colour444(height, width, seed, quality):
    return synthetic.encode_colour444(synthetic.colour_picture(height, width, seed), quality)


def same(a, b, label):
    assert a.result.iterations == b.result.iterations, label
    assert a.result.stop == b.result.stop, label
    assert np.array_equal(a.canvas, b.canvas), f"{label}: canvas"
    for ours, theirs in zip(a.coefficients, b.coefficients):
        assert np.array_equal(ours, theirs), f"{label}: coefficients"
    for name in ("primal", "dual"):
        assert np.array_equal(getattr(a.result.history, name), getattr(b.result.history, name)), f"{label}: {name}"
    assert np.array_equal(a.result.history.partial_gap, b.result.history.partial_gap, equal_nan=True), label


CASES = [
    ("tv grey", lambda: grey(120, 200, 51, 50), ["--method", "tv", "--iterations", "60", "--tolerance", "0"]),
    ("tgv grey odd", lambda: grey(104, 136, 52, 40), ["--method", "tgv", "--iterations", "60", "--tolerance", "0"]),
    ("tgv 4:2:0 free samples", lambda: colour(136, 200, 53, 75),
     ["--method", "tgv", "--iterations", "50", "--tolerance", "0"]),
    ("tgv 4:2:0 apart", lambda: colour(160, 96, 54, 60),
     ["--method", "tgv", "--channels", "apart", "--iterations", "40", "--tolerance", "0"]),
    ("tv 4:2:0 weights", lambda: colour(96, 320, 55, 80),
     ["--method", "tv", "--channel-weights", "1,0.5,0.75", "--iterations", "40", "--tolerance", "0"]),
    ("tgv 4:2:0 mixed", lambda: colour(144, 272, 56, 50),
     ["--method", "tgv", "--iterations", "40", "--tolerance", "0", "--precision", "mixed"]),
    ("tgv records every 7, partial gap", lambda: colour(112, 144, 57, 70),
     ["--method", "tgv", "--iterations", "45", "--record-every", "7", "--tolerance", "0", "--partial-radius", "20"]),
    ("tgv 4:4:4", lambda: colour444(88, 136, 60, 60), ["--method", "tgv", "--iterations", "40", "--tolerance", "0"]),
    ("tgv relative stop", lambda: grey(96, 104, 58, 50),
     ["--method", "tgv", "--iterations", "400", "--tolerance", "0", "--relative-tolerance", "3e-3"]),
]


@pytest.mark.parametrize("label,make,options", CASES, ids=[case[0] for case in CASES])
def test_sweep_equals_two_kernels(monkeypatch, label, make, options):
    image = make()
    print(f"sweep: {label} {image.height}x{image.width}", flush=True)
    monkeypatch.delenv("UNROUND_SWEEP", raising=False)
    apart = native.solve(image, options)
    monkeypatch.setenv("UNROUND_SWEEP", "1")
    fused = native.solve(image, options)
    same(fused, apart, label)


@pytest.mark.parametrize("label,make,options", CASES[1:4], ids=[case[0] for case in CASES[1:4]])
def test_sweep_does_not_depend_on_the_runs(monkeypatch, label, make, options):
    image = make()
    print(f"sweep runs: {label} {image.height}x{image.width}", flush=True)
    monkeypatch.setenv("UNROUND_SWEEP", "1")
    monkeypatch.setenv("UNROUND_DUAL_ROWS", "16")
    short = native.solve(image, options)
    monkeypatch.setenv("UNROUND_DUAL_ROWS", "64")
    long = native.solve(image, options)
    same(short, long, label)


def test_wide_planes_take_the_two_kernels(monkeypatch):
    # The sweep is built for buffer-resource planes only; wide planes fall back to the two
    # kernels, with the same values.
    image = colour(136, 200, 59, 75)
    options = ["--method", "tgv", "--iterations", "30", "--tolerance", "0"]
    print("sweep: wide planes", flush=True)
    monkeypatch.setenv("UNROUND_SWEEP", "1")
    narrow = native.solve(image, options)
    monkeypatch.setenv("UNROUND_WIDE_PLANES", "1")
    wide = native.solve(image, options)
    same(wide, narrow, "wide planes")
