"""The C interface without a GPU: the library loads, exports every function that
include/*.h declares, keeps the reference's layouts and ABI version, reads and writes the
settings as the reference does (rust/capi/tests/capi.rs:122-230), and refuses what it
should with the reference's statuses. No test here computes on a GPU."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import unround
from unround import native

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> list[str]:
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names += re.findall(r"\b(unround_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names) - {"unround_observer"})


def test_library_exports_every_declared_function():
    library = ctypes.CDLL(native.library_path())
    names = declared_functions()
    assert len(names) >= 35
    missing = [name for name in names if not hasattr(library, name)]
    assert missing == []


def test_versions_and_layouts():
    assert native._lib().unround_abi_version() == native.ABI_VERSION == 5
    assert "MI355X" in native.version()
    assert ctypes.sizeof(native.CRecord) == 64
    assert ctypes.sizeof(native.CComponent) == 152
    assert native.CComponent.quant_table.offset == 24
    assert ctypes.sizeof(native.CInput) == 24
    assert native.CInput.components.offset == 16


def test_settings_round_trip_through_the_library():
    given = unround.Settings(
        method="tv",
        data=(unround.DataTerm(mu=0.01), unround.DataTerm(mu=None, centres="mmse"), unround.DataTerm(slack=0.5)),
        tv=unround.TV(alpha=2.0, channel_weights=(1.0, 0.5, 0.25), coupled=False),
        pdhg=unround.PdhgOptions(iterations=123, tolerance=1e-7, step_ratio=30.0, relaxation=1.5, precision="double"),
    )
    held = native.held_options(given)
    assert "--method=tv" in held
    assert "--alpha=2.0" in held
    assert "--channel-weights=1.0,0.5,0.25" in held
    assert "--channels=apart" in held
    assert "--mu=0.01,rule,rule" in held
    assert "--centres=midpoint,mmse,midpoint" in held
    assert "--slack=0.0,0.0,0.5" in held
    assert "--iterations=123" in held
    assert "--tolerance=1e-7" in held
    assert "--precision=double" in held
    # What the library writes reads back as the same settings.
    assert native.held_options(held) == held


def test_numbers_read_back_to_the_last_bit():
    for value in (0.1, 1e-300, 123456789.123, 2.0**-40, 1 / 3):
        held = native.held_options(["--mu", repr(value)])
        written = next(option for option in held if option.startswith("--mu="))
        assert float(written.split("=", 1)[1]) == value


def test_every_refusal_is_named_at_once():
    with pytest.raises(native.UnroundError) as caught:
        native.held_options(["--mu", "-1", "--bits", "8", "--alpha=0", "--nope", "--relaxation", "2"])
    assert caught.value.status == native.Status.OPTIONS
    message = caught.value.message
    for part in ("--mu", "--bits", "--alpha", "--nope", "--relaxation"):
        assert part in message
    with pytest.raises(native.UnroundError):
        native.held_options(["--method", "tv", "--partial-radius", "3"])
    with pytest.raises(native.UnroundError):
        native.held_options(["--partial-tolerance", "1e-3"])
    with pytest.raises(native.UnroundError):
        native.held_options(["--mu", "1", "--mu", "2"])


def test_null_arguments_and_unsupported_paths_are_statuses():
    library = native._lib()
    message = ctypes.create_string_buffer(256)
    result = ctypes.c_void_p()
    callback = native.OBSERVER_TYPE()
    status = library.unround_solve(None, None, callback, None, ctypes.byref(result), message, 256)
    assert status == native.Status.ARGUMENT
    assert b"null" in message.value
    with pytest.raises(native.UnroundError) as caught:
        native.decode(b"\xff\xd8\xff", unround.Settings())
    assert caught.value.status == native.Status.READ
    image = unround.jpegio.make_image(
        8, 8, unround.jpegio.ColorSpace.GRAYSCALE,
        (unround.jpegio.make_component(np.zeros((1, 1, 8, 8), np.int16), np.ones(64, np.uint16)),))
    with pytest.raises(native.UnroundError) as caught:
        native.solve(image, unround.Settings(method="subgradient"))
    assert caught.value.status == native.Status.UNSUPPORTED
    # A message buffer too short is cut, NUL-terminated.
    short = ctypes.create_string_buffer(8)
    status = library.unround_solve(None, None, callback, None, ctypes.byref(result), short, 8)
    assert status == native.Status.ARGUMENT and len(short.value) == 7


def test_batch_refusals_come_before_the_gpu():
    library = native._lib()
    message = ctypes.create_string_buffer(256)
    results = (ctypes.c_void_p * 2)()
    status = library.unround_solve_batch(None, 2, None, results, None, message, 256)
    assert status == native.Status.ARGUMENT and b"null" in message.value
    # An empty batch is done at once; a bad input is named by its index.
    assert native.solve_batch([], ["--method", "tgv"]) == ([], [])
    grey = unround.jpegio.make_component(np.zeros((2, 2, 8, 8), np.int16), np.ones(64, np.uint16))
    good = unround.jpegio.make_image(16, 16, unround.jpegio.ColorSpace.GRAYSCALE, (grey,))
    bad = unround.jpegio.make_image(17, 16, unround.jpegio.ColorSpace.GRAYSCALE, (grey,))
    with pytest.raises(native.UnroundError) as caught:
        native.solve_batch([good, bad], ["--method", "tgv"])
    assert caught.value.status == native.Status.UNSUPPORTED and caught.value.message.startswith("input 1: ")
    with pytest.raises(native.UnroundError) as caught:
        native.solve_batch([good], ["--method", "mmse"])
    assert caught.value.status == native.Status.UNSUPPORTED


def test_samples_refusals_come_before_the_gpu():
    # What a PNG file's reconstruction refuses (rust/unround/src/rounded.rs:44-90;
    # rust/unround/src/samples.rs:74-110), refused before any GPU work.
    levels = np.zeros((4, 5, 3), np.uint16)
    cases = [
        (np.zeros((4, 5, 5), np.uint16), 8, ["--method", "tgv"], native.Status.ARGUMENT, "1 to 4 channels"),
        (levels, 3, ["--method", "tgv"], native.Status.OPTIONS, "bits"),
        (levels + 256, 8, ["--method", "tgv"], native.Status.OPTIONS, "at most 255"),
        (levels, 8, ["--method", "mmse"], native.Status.OPTIONS, "TV or TGV"),
        (levels, 8, ["--method", "tv", "--flat-weight", "1"], native.Status.OPTIONS, "--mu and --slack alone"),
        (levels, 8, ["--method", "tv", "--mu", "1,2,3"], native.Status.OPTIONS, "one value of each option"),
        (levels, 8, ["--method", "tv", "--channel-weights", "1,1"], native.Status.OPTIONS, "channel weights"),
    ]
    for given, bits, options, status, words in cases:
        with pytest.raises(native.UnroundError) as caught:
            native.solve_samples(given, bits, options)
        assert caught.value.status == status, (options, caught.value.message)
        assert words in caught.value.message, (options, caught.value.message)


def test_refused_frames():
    grey = unround.jpegio.make_component(np.zeros((2, 2, 8, 8), np.int16), np.ones(64, np.uint16))
    # Blocks that do not cover the picture, a colour space with the wrong count of
    # components, a step of 0, channel weights not one per channel.
    cases = [
        (unround.jpegio.make_image(17, 16, unround.jpegio.ColorSpace.GRAYSCALE, (grey,)), unround.Settings(), 3),
        (unround.jpegio.make_image(16, 16, unround.jpegio.ColorSpace.YCBCR, (grey,)), unround.Settings(), 3),
        (unround.jpegio.make_image(16, 16, unround.jpegio.ColorSpace.GRAYSCALE,
                              (unround.jpegio.make_component(np.zeros((2, 2, 8, 8), np.int16), np.zeros(64, np.uint16)),)),
         unround.Settings(), 4),
        (unround.jpegio.make_image(16, 16, unround.jpegio.ColorSpace.GRAYSCALE, (grey,)),
         unround.Settings(tgv=unround.TGV(channel_weights=(1.0, 2.0))), 4),
    ]
    for image, settings, status in cases:
        with pytest.raises(native.UnroundError) as caught:
            native.solve(image, settings)
        assert caught.value.status == status, caught.value.message


def test_jfif_conversions_invert_each_other():
    rng = np.random.default_rng(1)
    picture = rng.uniform(0, 255, size=(5, 7, 3))
    planes = unround.to_ycbcr(picture)
    assert np.max(np.abs(unround.to_rgb(planes) - picture)) < 1e-12
    # The exact inverse of docs/math.md:1004-1005 at a grey pixel.
    grey = unround.to_rgb(np.array([[[100.0]], [[128.0]], [[128.0]]]))
    assert np.all(grey == 100.0)


@pytest.mark.parametrize("height,width", [(1 << 33, 1 << 33), (1 << 62, 4), ((1 << 64) - 1, (1 << 64) - 1),
                                          (1 << 31, 1 << 31)])
def test_jfif_conversions_refuse_sizes_that_overflow(height, width):
    # 3 x height x width doubles beyond what a size can hold (or a pointer difference can
    # span) is refused before any sample is read, as the reference's conversions refuse it
    # (rust/capi/src/lib.rs:1568-1582): the one-double buffers here are never touched.
    lib = native._lib()
    room = (ctypes.c_double * 1)()
    out = (ctypes.c_double * 1)()
    for name in ("unround_to_rgb", "unround_to_ycbcr"):
        function = getattr(lib, name)
        status = function(ctypes.cast(room, ctypes.POINTER(ctypes.c_double)), ctypes.c_uint64(height),
                          ctypes.c_uint64(width), ctypes.cast(out, ctypes.POINTER(ctypes.c_double)))
        assert status == native.Status.ARGUMENT, name
