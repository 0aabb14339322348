"""The GPU's binary32 path against the oracle (binary64 CPU) on the same seeded synthetic
inputs, at sizes the oracle finishes in seconds: 8-bit outputs within 1 level, differing on
at most 0.1 % of the pixels; PSNR against the uncompressed picture within 0.01 dB; every
coefficient of the result requantizing to its level, bit for bit; the records within
binary32's reach of the oracle's. Run with -s to see the measured distances."""

import numpy as np
import pytest

import helpers
from unround import native, synthetic

pytestmark = pytest.mark.gpu


check = helpers.parity_check


def test_tv_grey_256_q50_200_iterations():
    picture = synthetic.grey_picture(256, 256, 0)
    image = synthetic.encode_grey(picture, 50)
    check(image, picture, ["--method", "tv", "--iterations", "200", "--tolerance", "0"])


def test_tgv_grey_odd_size():
    picture = synthetic.grey_picture(120, 200, 3)
    image = synthetic.encode_grey(picture, 40)
    check(image, picture, ["--method", "tgv", "--iterations", "150", "--tolerance", "0"])


def test_tgv_colour420_free_samples():
    # 136 x 200 is not a multiple of 16: samples beyond the blocks are free.
    picture = synthetic.colour_picture(136, 200, 5)
    image = synthetic.encode_colour420(picture, 75)
    check(image, picture, ["--method", "tgv", "--iterations", "100", "--tolerance", "0"])


def test_tv_colour420_apart_weights():
    picture = synthetic.colour_picture(96, 128, 6)
    image = synthetic.encode_colour420(picture, 60)
    check(image, picture, ["--method", "tv", "--channels", "apart", "--channel-weights", "1,0.5,0.75", "--iterations",
                           "80", "--tolerance", "0"])


def test_tgv_grey_partial_gap_and_model_options():
    picture = synthetic.grey_picture(64, 96, 8)
    image = synthetic.encode_grey(picture, 30)
    options = ["--method", "tgv", "--iterations", "60", "--tolerance", "0", "--partial-radius", "20", "--slack",
               "0.5", "--slack-cost", "3", "--centres", "mmse", "--flat-weight", "0.3"]
    decoded, expected = check(image, picture, options)
    ours = decoded.result.history.partial_gap
    assert np.max(np.abs(ours - expected.partial) / np.abs(expected.partial)) < 1e-4


def test_tolerance_stop_matches():
    picture = synthetic.grey_picture(64, 64, 9)
    image = synthetic.encode_grey(picture, 50)
    options = ["--method", "tv", "--iterations", "2000"]
    expected = helpers.oracle_solve(image, options)
    decoded = native.solve(image, options)
    assert expected.stop == 1
    assert decoded.result.stop == native.Stop.CONVERGED
    # binary32 may reach the tolerance a record apart from binary64.
    assert abs(decoded.result.iterations - expected.iterations) <= 10


def test_mixed_precision_bf16_duals_error_against_binary64():
    # p and r stored in bfloat16 (8-bit significands), binary32 arithmetic: the error
    # against the binary64 oracle is reported; the requantization stays exact.
    picture = synthetic.grey_picture(256, 256, 11)
    image = synthetic.encode_grey(picture, 40)
    options = ["--method", "tgv", "--iterations", "200", "--tolerance", "0"]
    expected = helpers.oracle_solve(image, options)
    decoded = native.solve(image, options + ["--precision", "mixed"])
    worst = float(np.max(np.abs(decoded.planes - expected.canvas[:, :256, :256])))
    ours, theirs = helpers.psnr(decoded.planes[0], picture), helpers.psnr(expected.canvas[0], picture)
    print(f"  mixed: max |diff| {worst:.3g}, PSNR {ours:.4f} vs {theirs:.4f}")
    assert worst < 0.5
    assert abs(ours - theirs) < 0.01
    levels = helpers.requantized(decoded.canvas[0], image, 0, (1, 1), 32, 32)
    assert np.array_equal(levels, helpers.kept_levels(image, 0, 32, 32))


def test_runs_repeat_to_the_last_bit():
    picture = synthetic.colour_picture(64, 80, 12)
    image = synthetic.encode_colour420(picture, 50)
    options = ["--method", "tgv", "--iterations", "40", "--tolerance", "0"]
    first, second = native.solve(image, options), native.solve(image, options)
    assert np.array_equal(first.canvas, second.canvas)
    assert np.array_equal(first.result.history.primal, second.result.history.primal)
    assert np.array_equal(first.result.history.dual, second.result.history.dual)


@pytest.mark.parametrize("mode", ["1", "2"])
def test_timing_leaves_the_warm_up_untimed_and_the_result_unchanged(monkeypatch, mode):
    picture = synthetic.colour_picture(64, 80, 13)
    image = synthetic.encode_colour420(picture, 50)
    options = ["--method", "tgv", "--iterations", "12", "--tolerance", "0", "--record-every", "0"]
    plain = native.solve(image, options)
    assert plain.kernel_times is None
    monkeypatch.setenv("UNROUND_TIMING", mode)
    monkeypatch.setenv("UNROUND_TIMING_WARMUP", "5")
    timed = native.solve(image, options)
    times = timed.kernel_times
    assert times.iterations == 7 and times.iterations_seconds > 0
    if mode == "2":
        # Y, then Cb and Cr in one launch: two proximal launches an iteration.
        assert times.dual_launches == 7 and times.primal_launches == 14
    else:
        assert times.dual_launches == 0 and times.primal_launches == 0
    assert np.array_equal(timed.canvas, plain.canvas)


def encode_colour411(picture: np.ndarray, quality: int):
    """An RGB picture as YCbCr 4:1:1 (luma sampled 4 across and 1 down, chroma cells 4
    samples wide and 1 high), quantized at a quality with the scaled tables."""
    from unround import jpegio

    height, width, _ = picture.shape
    planes = synthetic._ycbcr_planes(picture)
    rows, columns = -(-height // 8) * 8, -(-width // 32) * 32
    luma = synthetic._padded(planes[0], rows, columns)
    chroma = [np.floor(synthetic._padded(p, rows, columns).reshape(rows, columns // 4, 4).mean(axis=2) + 0.5)
              for p in planes[1:]]
    luma_table = synthetic.ijg_table(synthetic.LUMA, quality)
    chroma_table = synthetic.ijg_table(synthetic.CHROMA, quality)
    components = (jpegio.make_component(synthetic._levels(luma, luma_table), luma_table, 4, 1),
                  jpegio.make_component(synthetic._levels(chroma[0], chroma_table), chroma_table, 1, 1),
                  jpegio.make_component(synthetic._levels(chroma[1], chroma_table), chroma_table, 1, 1))
    return jpegio.make_image(height, width, jpegio.ColorSpace.YCBCR, components)


def test_tgv_colour411_cells_of_any_shape():
    # Chroma cells 4 x 1 take the prox and the conjugates' instances for any cell shape;
    # the dual values of the records (the conjugates' sums) match the oracle's too.
    picture = synthetic.colour_picture(72, 136, 14)
    image = encode_colour411(picture, 70)
    options = ["--method", "tgv", "--iterations", "40", "--tolerance", "0", "--record-every", "8"]
    decoded, expected = check(image, picture, options)
    history = decoded.result.history
    relative = np.abs(history.dual - expected.dual) / np.abs(expected.primal)
    print(f"  records: dual rel {np.max(relative):.3g}")
    assert np.max(relative) < 1e-4
