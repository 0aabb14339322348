"""Many images solved together (unround_solve_batch) against each image solved alone: the
batch runs every image's tiles in shared launches, with the records and stopping tests on
the device and each run of iterations between records replayed from a HIP graph, yet each
image's canvas, coefficients, records and stopping iteration must be its own solve's, bit
for bit. Sizes not multiples of 16 give 4:2:0 frames free samples; a relative tolerance
stops images at different records, so that stopped images sit out later graph replays."""

import numpy as np
import pytest

from unround import native, synthetic

pytestmark = pytest.mark.gpu


def mixed_images(specs):
    images = []
    for seed, (height, width, colour, quality) in enumerate(specs, start=1):
        if colour:
            images.append(synthetic.encode_colour420(synthetic.colour_picture(height, width, seed), quality))
        else:
            images.append(synthetic.encode_grey(synthetic.grey_picture(height, width, seed), quality))
    return images


def same_as_alone(images, options):
    print(f"batch of {len(images)}: {' '.join(options)}", flush=True)
    together, latencies = native.solve_batch(images, options)
    assert len(together) == len(images) and len(latencies) == len(images)
    for index, (image, batched) in enumerate(zip(images, together)):
        alone = native.solve(image, options)
        print(f"  image {index} {image.height}x{image.width} space {int(image.color_space)}: "
              f"{alone.result.iterations} iterations, stop {alone.result.stop.name}, latency {latencies[index]:.4f} s",
              flush=True)
        assert batched.result.iterations == alone.result.iterations
        assert batched.result.stop == alone.result.stop
        assert np.array_equal(batched.canvas, alone.canvas), f"image {index}: canvas"
        for ours, theirs in zip(batched.coefficients, alone.coefficients):
            assert np.array_equal(ours, theirs), f"image {index}: coefficients"
        a, b = batched.result.history, alone.result.history
        assert list(a.iterations) == list(b.iterations)
        for name in ("primal", "dual"):
            assert np.array_equal(getattr(a, name), getattr(b, name)), f"image {index}: {name} records"
        assert np.array_equal(a.scaling, b.scaling, equal_nan=True)
        assert np.array_equal(a.partial_gap, b.partial_gap, equal_nan=True)
    return together, latencies


SPECS = [(96, 104, False, 50), (120, 88, True, 75), (64, 64, False, 30), (136, 200, True, 90), (72, 256, False, 60)]


def test_tgv_batch_is_each_image_alone():
    images = mixed_images(SPECS)
    together, latencies = same_as_alone(
        images, ["--method", "tgv", "--iterations", "300", "--relative-tolerance", "2e-3", "--tolerance", "0"])
    stops = {decoded.result.iterations for decoded in together}
    assert len(stops) > 1, "the images stop at different records"
    assert all(value >= 0.0 for value in latencies)


def test_tv_batch_is_each_image_alone():
    same_as_alone(mixed_images(SPECS[:3]), ["--method", "tv", "--iterations", "120", "--tolerance", "0",
                                            "--relative-tolerance", "1e-3"])


def test_mixed_precision_batch_is_each_image_alone():
    same_as_alone(mixed_images(SPECS[1:4]), ["--method", "tgv", "--iterations", "60", "--tolerance", "0",
                                             "--precision", "mixed"])


def test_records_every_seven_alternate_parity():
    # Runs of 7 iterations end on the other buffer: the graphs of both parities replay.
    same_as_alone(mixed_images(SPECS[:2]), ["--method", "tgv", "--iterations", "50", "--record-every", "7",
                                            "--tolerance", "0"])


def test_no_records_until_the_end():
    same_as_alone(mixed_images(SPECS[2:4]), ["--method", "tgv", "--iterations", "35", "--record-every", "0",
                                             "--tolerance", "0"])


def test_long_runs_with_short_tails(monkeypatch):
    # Runs of 32 rows: each XCD's share of the dual step's table holds long runs first and
    # its last stretch cut into runs of 8 (the shares here are a few times that stretch),
    # with pictures whose widths leave strip groups part empty.
    monkeypatch.setenv("UNROUND_DUAL_ROWS", "32")
    images = mixed_images([(2048, 2040, True, 75), (1024, 1544, False, 60), (200, 136, True, 40)])
    same_as_alone(images, ["--method", "tgv", "--iterations", "20", "--tolerance", "0"])
