"""unround_solve_scratch, the solve that allocates no device memory (its buffers are
stretches of a fixed region the library holds on each GPU), against unround_solve of the
same input: the planes bit for bit, the iterations and the stop reason. Two calls in a row
share the region, the second with a smaller input, so that the first call's leftovers in
the region must not reach the second; inputs too large for the region, and planes of the
wrong size, are refused before anything runs."""

import ctypes

import numpy as np
import pytest

from unround import native, synthetic

pytestmark = pytest.mark.gpu


def scratch_solve(image, options):
    library = native._lib()
    held = native.SettingsHandle(options)
    marshalled = native.Marshalled(image)
    planes = np.full((len(image.components), int(image.height), int(image.width)), np.nan)
    iterations, stop = ctypes.c_uint64(), ctypes.c_int32()
    room = ctypes.create_string_buffer(native.MESSAGE_BYTES)
    status = library.unround_solve_scratch(ctypes.byref(marshalled.input), held.handle,
                                           planes.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), planes.size,
                                           ctypes.byref(iterations), ctypes.byref(stop), room, native.MESSAGE_BYTES)
    if status != native.Status.OK:
        raise native.status_error(int(status), room.value.decode())
    return planes, int(iterations.value), int(stop.value)


def same_as_solve(image, options, label):
    print(f"scratch: {label} {image.height}x{image.width} {' '.join(options)}", flush=True)
    planes, iterations, stop = scratch_solve(image, options)
    alone = native.solve(image, options)
    assert np.array_equal(planes, alone.planes), label
    assert iterations == alone.result.iterations and stop == int(alone.result.stop), label


def test_two_calls_in_a_row_the_second_smaller():
    first = synthetic.encode_colour420(synthetic.colour_picture(200, 136, 41), 75)
    second = synthetic.encode_grey(synthetic.grey_picture(96, 80, 42), 50)
    same_as_solve(first, ["--method", "tgv", "--iterations", "60", "--tolerance", "0"], "first, 4:2:0 TGV")
    same_as_solve(second, ["--method", "tv", "--iterations", "40", "--tolerance", "0"], "second, grey TV")


def test_precisions_and_a_tolerance_stop():
    image = synthetic.encode_grey(synthetic.grey_picture(64, 64, 9), 50)
    same_as_solve(image, ["--method", "tv", "--iterations", "2000"], "stops by its tolerance")
    colour = synthetic.encode_colour420(synthetic.colour_picture(64, 80, 43), 60)
    same_as_solve(colour, ["--method", "tgv", "--iterations", "30", "--tolerance", "0", "--precision", "mixed"], "mixed")
    same_as_solve(image, ["--method", "tgv", "--iterations", "30", "--tolerance", "0", "--precision", "double"],
                  "double")


def test_the_largest_grey_picture_the_docs_name_fits():
    image = synthetic.encode_grey(synthetic.grey_picture(512, 512, 44), 30)
    same_as_solve(image, ["--method", "tgv", "--iterations", "20", "--tolerance", "0"], "512x512 grey")


def test_refusals_come_before_the_gpu():
    large = synthetic.encode_grey(synthetic.grey_picture(1024, 1024, 45), 50)
    with pytest.raises(native.UnroundError) as caught:
        scratch_solve(large, ["--method", "tgv", "--iterations", "10"])
    assert caught.value.status == native.Status.ARGUMENT and "bytes of device memory" in caught.value.message
    small = synthetic.encode_grey(synthetic.grey_picture(16, 16, 46), 50)
    library = native._lib()
    held = native.SettingsHandle(["--method", "tv"])
    marshalled = native.Marshalled(small)
    planes = np.zeros(10)
    room = ctypes.create_string_buffer(256)
    status = library.unround_solve_scratch(ctypes.byref(marshalled.input), held.handle,
                                           planes.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), planes.size, None,
                                           None, room, 256)
    assert status == native.Status.ARGUMENT and b"not 10" in room.value
    # And the region still serves a solve after the refusals.
    same_as_solve(small, ["--method", "tv", "--iterations", "10", "--tolerance", "0"], "after refusals")
