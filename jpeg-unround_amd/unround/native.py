"""JPEG-Unround on an AMD Instinct GPU, through its C interface (include/unround.h).

The same functions, types and errors as the reference's python/src/unround/native.py for
the reconstruction of components given as arrays: options() and held_options() turn the
settings into the command line's options, solve() reconstructs an Image on the GPU, and
its Decoded result, records and errors read as the reference's. The library is the one at
the path in UNROUND_LIBRARY, or else unround/_native/libunround_capi.so, which
jpeg-unround_amd/build.py writes; a missing library raises LibraryNotFoundError, and no
computation falls back to the CPU. ctypes releases the GIL while the library computes.
"""

from __future__ import annotations

import ctypes
import enum
import functools
import os
import warnings
from collections.abc import Callable
from dataclasses import dataclass
from pathlib import Path
from typing import Final, Sequence

import numpy as np
import numpy.typing as npt

from unround import _loading, jpegio
from unround.settings import TGV, TV, DataTerm, PdhgOptions, Settings, SubgradientOptions

__all__ = [
    "ABI_VERSION",
    "Decoded",
    "History",
    "KernelTimes",
    "LibraryNotFoundError",
    "Problem",
    "Record",
    "Solution",
    "Status",
    "Stop",
    "UnroundError",
    "UnroundWarning",
    "decode",
    "held_options",
    "options",
    "oriented",
    "png_bytes",
    "pnm_bytes",
    "project",
    "solve",
    "solve_batch",
    "solve_samples",
    "to_rgb",
    "tiff_bytes",
    "to_ycbcr",
    "version",
]

Array = npt.NDArray[np.float64]
Observer = Callable[["Record"], "bool | None"]

ABI_VERSION: Final = 5

_MESSAGE_SIZE: Final = 4096
_FORMAT_PNG: Final = 2
_BLOCK_SIZE: Final = 64


class _Record(ctypes.Structure):
    _fields_ = (
        ("iteration", ctypes.c_uint64),
        ("gap", ctypes.c_double),
        ("primal", ctypes.c_double),
        ("dual", ctypes.c_double),
        ("channels", ctypes.c_uint64),
        ("height", ctypes.c_uint64),
        ("width", ctypes.c_uint64),
        ("canvas", ctypes.POINTER(ctypes.c_double)),
    )


class _Component(ctypes.Structure):
    _fields_ = (
        ("levels", ctypes.POINTER(ctypes.c_int16)),
        ("rows", ctypes.c_uint32),
        ("columns", ctypes.c_uint32),
        ("h_samp_factor", ctypes.c_int32),
        ("v_samp_factor", ctypes.c_int32),
        ("quant_table", ctypes.c_uint16 * _BLOCK_SIZE),
    )


class _Input(ctypes.Structure):
    _fields_ = (
        ("height", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("color_space", ctypes.c_int32),
        ("count", ctypes.c_uint32),
        ("components", ctypes.POINTER(_Component)),
    )


class _Start(ctypes.Structure):
    _fields_ = (
        ("coefficients", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
        ("p", ctypes.POINTER(ctypes.c_double)),
        ("w", ctypes.POINTER(ctypes.c_double)),
        ("r", ctypes.POINTER(ctypes.c_double)),
    )


_SIZES: Final = ((_Record, 64), (_Component, 152), (_Input, 24))

_OBSERVER: Final = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(_Record))


class Status(enum.IntEnum):
    """What a function of the interface reports: unround_status."""

    OK = 0
    ARGUMENT = 1
    READ = 2
    UNSUPPORTED = 3
    OPTIONS = 4
    IO = 5
    INTERNAL = 6


class Stop(enum.IntEnum):
    """Why a solver stopped."""

    CONVERGED = 1
    ITERATIONS = 2
    OBSERVER = 3
    STATIONARY = 4


class UnroundError(Exception):
    """Arrays that were not reconstructed: the status of the interface, and its message."""

    def __init__(self, status: Status, message: str) -> None:
        super().__init__(f"{message} ({status.name.lower()})")
        self.status = status
        self.message = message

    def __reduce__(self) -> tuple[type["UnroundError"], tuple[Status, str]]:
        return (type(self), (self.status, self.message))


class LibraryNotFoundError(RuntimeError):
    """The library is not where this module looks for it, or is not of its version."""


class UnroundWarning(UserWarning):
    """A warning of the library."""


def _declare(library: ctypes.CDLL, name: str, result: object, *arguments: object) -> None:
    function = getattr(library, name)
    function.argtypes = arguments
    function.restype = result


def _declare_all(library: ctypes.CDLL) -> None:
    size, handle, u64, i32 = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32
    doubles, u64_out, i32_out = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64), ctypes.POINTER(i32)
    strings = ctypes.POINTER(ctypes.c_char_p)
    message = (ctypes.c_char_p, size)
    made = (_OBSERVER, handle, ctypes.POINTER(handle), *message)
    _declare(library, "unround_abi_version", i32)
    _declare(library, "unround_version", ctypes.c_char_p)
    _declare(library, "unround_settings_new", handle, size, strings, *message)
    _declare(library, "unround_settings_free", None, handle)
    _declare(library, "unround_settings_options", size, handle, ctypes.c_char_p, size)
    _declare(library, "unround_decode", i32, ctypes.c_char_p, size, handle, *made)
    _declare(library, "unround_solve", i32, ctypes.POINTER(_Input), handle, *made)
    _declare(library, "unround_solve_start", i32, ctypes.POINTER(_Input), handle, ctypes.POINTER(_Start), *made)
    _declare(library, "unround_solve_batch", i32, ctypes.POINTER(_Input), size, handle, ctypes.POINTER(handle), doubles,
             *message)
    _declare(library, "unround_solve_samples", i32, ctypes.POINTER(ctypes.c_uint16), u64, u64, u64, ctypes.c_uint32,
             handle, *made)
    _declare(library, "unround_result_alpha_solver", i32, handle, u64_out, i32_out)
    _declare(library, "unround_result_alpha_record_iterations", u64_out, handle, u64_out)
    _declare(library, "unround_result_alpha_record_values", doubles, handle, i32, u64_out)
    _declare(library, "unround_result_free", None, handle)
    _declare(library, "unround_result_picture", doubles, handle, u64_out, u64_out, u64_out)
    _declare(library, "unround_result_planes", doubles, handle, u64_out)
    _declare(library, "unround_result_color_space", i32, handle)
    _declare(library, "unround_result_canvas", doubles, handle, u64_out, u64_out, u64_out)
    _declare(library, "unround_result_components", u64, handle)
    _declare(library, "unround_result_coefficients", doubles, handle, u64, u64_out)
    _declare(library, "unround_result_problem", doubles, handle, u64, i32, u64_out)
    _declare(library, "unround_result_component", i32, handle, u64, u64_out, u64_out, u64_out, u64_out)
    _declare(library, "unround_result_solver", i32, handle, u64_out, i32_out)
    _declare(library, "unround_result_record_iterations", u64_out, handle, u64_out)
    _declare(library, "unround_result_record_values", doubles, handle, i32, u64_out)
    _declare(library, "unround_result_kernel_times", i32, handle, doubles, size)
    _declare(library, "unround_result_exif_orientation", i32, handle)
    _declare(library, "unround_result_icc_profile", ctypes.POINTER(ctypes.c_uint8), handle, u64_out)
    _declare(library, "unround_result_excess", doubles, handle, u64, u64_out)
    _declare(library, "unround_project", i32, ctypes.c_char_p, size, handle, doubles, u64, u64, u64,
             ctypes.POINTER(handle), *message)
    _declare(library, "unround_result_format", i32, handle)
    _declare(library, "unround_result_bits", i32, handle)
    _declare(library, "unround_result_warning", ctypes.c_char_p, handle)
    bytes_out = (ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)), u64_out, *message)
    icc = (ctypes.c_char_p, u64)
    _declare(library, "unround_tiff", i32, doubles, u64, u64, u64, i32, *icc, *bytes_out)
    _declare(library, "unround_png", i32, doubles, u64, u64, u64, i32, i32, i32, *icc, *bytes_out)
    _declare(library, "unround_pnm", i32, doubles, u64, u64, u64, i32, i32, *bytes_out)
    _declare(library, "unround_bytes_free", None, ctypes.POINTER(ctypes.c_uint8), u64)
    _declare(library, "unround_orient", i32, doubles, u64, u64, u64, i32, doubles, u64_out, u64_out, *message)
    _declare(library, "unround_to_rgb", i32, doubles, u64, u64, doubles)
    _declare(library, "unround_to_ycbcr", i32, doubles, u64, u64, doubles)


@functools.cache
def _library() -> ctypes.CDLL:
    for structure, wanted in _SIZES:
        if ctypes.sizeof(structure) != wanted:
            raise RuntimeError(f"{structure.__name__} is {ctypes.sizeof(structure)} bytes, and the interface's is {wanted}")
    path = _loading.configured("UNROUND_LIBRARY") or _loading.bundled(_loading.LIBRARY)
    if path is None or not path.is_file():
        where = f"UNROUND_LIBRARY names {path}, where there is no file" if path else "the package holds no library"
        raise LibraryNotFoundError(f"{where}: build it with python jpeg-unround_amd/build.py")
    library = ctypes.CDLL(str(path))
    _declare_all(library)
    found = int(library.unround_abi_version())
    if found != ABI_VERSION:
        raise LibraryNotFoundError(f"{path} is of ABI version {found}, and this module of version {ABI_VERSION}")
    return library


def library_path() -> str:
    """The path of the library this module loaded."""
    return str(_library()._name)


def version() -> str:
    """The version of JPEG-Unround and of the implementation, as the library says them."""
    return _library().unround_version().decode("utf-8")


def _number(value: float) -> str:
    return repr(float(value))


def _data_options(data: DataTerm | tuple[DataTerm, ...]) -> list[str]:
    """The options of G: one value where every component has it, or one for each
    (python/src/unround/native.py:245-271)."""
    terms = (data,) if isinstance(data, DataTerm) else data
    fields: tuple[tuple[str, Callable[[DataTerm], str]], ...] = (
        ("--mu", lambda term: "rule" if term.mu is None else _number(term.mu)),
        ("--mu-scale", lambda term: _number(term.mu_scale)),
        ("--mu-power", lambda term: _number(term.mu_power)),
        ("--mu-chroma", lambda term: _number(term.mu_chroma)),
        ("--weight-power", lambda term: _number(term.power)),
        ("--dc-weight", lambda term: _number(term.dc_weight)),
        ("--centres", lambda term: term.centres),
        ("--slack", lambda term: _number(term.slack)),
        ("--slack-cost", lambda term: _number(term.slack_cost)),
        ("--flat-weight", lambda term: _number(term.flat_weight)),
        ("--flat-levels", lambda term: str(term.flat_levels)),
        ("--flat-dc", lambda term: "any" if term.flat_dc is None else str(term.flat_dc)),
    )
    result = []
    for name, value in fields:
        values = [value(term) for term in terms]
        result += [name, values[0] if len(set(values)) == 1 else ",".join(values)]
    return result


def _pdhg_options(solver: PdhgOptions) -> list[str]:
    result = [
        "--relative-tolerance", _number(solver.relative_tolerance),
        "--partial-tolerance", _number(solver.partial_tolerance),
        "--step-product", _number(solver.step_product),
        "--record-every", str(solver.record_every),
        "--free-radius", _number(solver.free_radius),
        "--threads", str(solver.threads),
        "--precision", solver.precision,
    ]
    if solver.iterations is not None:
        result += ["--iterations", str(solver.iterations)]
    for name, value in (
        ("--tolerance", solver.tolerance),
        ("--step-ratio", solver.step_ratio),
        ("--relaxation", solver.relaxation),
        ("--norm-squared", solver.norm_squared),
        ("--partial-radius", solver.partial_radius),
    ):
        if value is not None:
            result += [name, _number(value)]
    if not solver.scale_with_weight:
        result.append("--no-weight-scaling")
    return result


def _subgradient_options(method: SubgradientOptions) -> list[str]:
    result = [
        "--subgradient-iterations", str(method.iterations),
        "--subgradient-record-every", str(method.record_every),
        "--subgradient-step", _number(method.step),
        "--subgradient-decay", _number(method.decay),
    ]
    if not method.momentum:
        result.append("--no-momentum")
    return result


def options(settings: Settings | None = None, *, max_pixels: int = 0, max_scans: int = 0,
            warnings_are_errors: bool = False) -> list[str]:
    """The command line's options that are the settings (python/src/unround/native.py:321-353),
    with the limits of reading a file."""
    reading = ([f"--max-pixels={max_pixels}"] if max_pixels else []) + (
        [f"--max-scans={max_scans}"] if max_scans else []) + (["--warnings-are-errors"] if warnings_are_errors else [])
    if isinstance(settings, list):
        return settings + reading
    settings = Settings() if settings is None else settings
    weights: TV | TGV = settings.tgv if settings.method == "tgv" else settings.tv
    result = ["--method", settings.method, "--alpha", _number(settings.tv.alpha)]
    result += ["--alpha1", _number(settings.tgv.alpha1), "--alpha0", _number(settings.tgv.alpha0)]
    if weights.channel_weights is not None:
        result += ["--channel-weights", ",".join(_number(value) for value in weights.channel_weights)]
    result += ["--channels", "coupled" if weights.coupled else "apart"]
    result += _data_options(settings.data)
    result += _pdhg_options(settings.pdhg)
    result += _subgradient_options(settings.subgradient)
    return result + reading


class _Settings:
    """Settings of the interface, made from options, and freed with the object."""

    def __init__(self, given: list[str]) -> None:
        self._library = _library()
        encoded = [option.encode("utf-8") for option in given]
        array = (ctypes.c_char_p * max(len(encoded), 1))(*encoded)
        message = ctypes.create_string_buffer(_MESSAGE_SIZE)
        self.handle: int | None = self._library.unround_settings_new(len(encoded), array, message, _MESSAGE_SIZE)
        if not self.handle:
            raise UnroundError(Status.OPTIONS, message.value.decode("utf-8", errors="replace"))

    def __del__(self) -> None:
        handle = getattr(self, "handle", None)
        if handle:
            self._library.unround_settings_free(handle)
            self.handle = None


def held_options(settings: Settings | list[str] | None = None) -> list[str]:
    """The settings as the library holds them: every value, as --name=value options."""
    library = _library()
    handle = _Settings(settings if isinstance(settings, list) else options(settings))
    length = int(library.unround_settings_options(handle.handle, None, 0))
    buffer = ctypes.create_string_buffer(length + 1)
    library.unround_settings_options(handle.handle, buffer, length + 1)
    return buffer.value.decode("ascii").split(" ")


@dataclass(frozen=True, slots=True, eq=False)
class Record:
    """A record of the solver, as an observer sees it (gap per sample; canvas (C, H, W))."""

    iteration: int
    gap: float
    primal: float
    dual: float
    canvas: Array


@dataclass(frozen=True, slots=True, eq=False)
class Problem:
    """A component's model: cell ratio, interval ends and centres (rows, columns, 8, 8), steps and weights (8, 8)."""

    ratio: tuple[int, int]
    lower: Array
    upper: Array
    centres: Array
    steps: Array
    weights: Array


@dataclass(frozen=True, slots=True, eq=False)
class History:
    """The solver's records: iterations, seconds, primal and dual values, TGV's scaling and partial gap."""

    iterations: npt.NDArray[np.int64]
    seconds: Array
    primal: Array
    dual: Array
    scaling: Array
    partial_gap: Array

    @property
    def gap(self) -> Array:
        return np.asarray(self.primal - self.dual, dtype=np.float64)


@dataclass(frozen=True, slots=True, eq=False)
class Solution:
    """What the solver did: its iterations, why it stopped, and its records."""

    iterations: int
    stop: Stop
    history: History

    @property
    def converged(self) -> bool:
        return self.stop is Stop.CONVERGED


@dataclass(frozen=True, slots=True)
class KernelTimes:
    """HIP-event durations of a solve's kernels: UNROUND_TIMING=1 during the solve times the
    runs of iterations between records only, UNROUND_TIMING=2 also every launch."""

    primal_seconds: float
    primal_launches: int
    dual_seconds: float
    dual_launches: int
    iterations_seconds: float
    iterations: int


@dataclass(frozen=True, slots=True, eq=False)
class Decoded:
    """Arrays reconstructed on the GPU, as the reference's Decoded (python/src/unround/native.py:520-560)."""

    picture: Array
    planes: Array
    canvas: Array
    color_space: jpegio.ColorSpace
    coefficients: tuple[Array, ...]
    problems: tuple[Problem, ...]
    result: Solution | None
    icc_profile: bytes | None
    exif_orientation: int
    input_format: str = "jpeg"
    bits: int = 8
    alpha_result: Solution | None = None
    excess: tuple[Array, ...] | None = None
    kernel_times: KernelTimes | None = None


def _copy(pointer: object, shape: tuple[int, ...]) -> Array:
    count = int(np.prod(shape, dtype=np.int64))
    if count == 0:
        return np.zeros(shape, dtype=np.float64)
    if not pointer:
        raise RuntimeError("the library gave no values where it has some")
    values = np.ctypeslib.as_array(pointer, shape=(count,))
    return np.array(values, dtype=np.float64, copy=True).reshape(shape)


def _checked(pointer: object, count: ctypes.c_uint64, shape: tuple[int, ...]) -> Array:
    if int(count.value) != int(np.prod(shape, dtype=np.int64)):
        raise RuntimeError(f"the library gave {count.value} values where {shape} holds {int(np.prod(shape))}")
    return _copy(pointer, shape)


def _problem(library: ctypes.CDLL, result: ctypes.c_void_p, index: int) -> tuple[Array, Problem]:
    rows, columns, down, across = (ctypes.c_uint64() for _ in range(4))
    status = library.unround_result_component(result, index, ctypes.byref(rows), ctypes.byref(columns),
                                              ctypes.byref(down), ctypes.byref(across))
    if status != Status.OK:
        raise RuntimeError(f"the result has no component {index}")
    shape = (int(rows.value), int(columns.value), 8, 8)
    count = ctypes.c_uint64()
    coefficients = _checked(library.unround_result_coefficients(result, index, ctypes.byref(count)), count, shape)
    arrays = []
    for which, wanted in ((0, shape), (1, shape), (2, shape), (3, (8, 8)), (4, (8, 8))):
        arrays.append(_checked(library.unround_result_problem(result, index, which, ctypes.byref(count)), count, wanted))
    lower, upper, centres, steps, weights = arrays
    return coefficients, Problem((int(down.value), int(across.value)), lower, upper, centres, steps, weights)


def _solution(library: ctypes.CDLL, result: ctypes.c_void_p, *, alpha: bool = False) -> Solution | None:
    """The solver of the result, or of its alpha (python/src/unround/native.py:548-575)."""
    solver, recorded_at, values_of = (
        (library.unround_result_alpha_solver, library.unround_result_alpha_record_iterations,
         library.unround_result_alpha_record_values) if alpha else
        (library.unround_result_solver, library.unround_result_record_iterations, library.unround_result_record_values))
    iterations, stop = ctypes.c_uint64(), ctypes.c_int32()
    if not solver(result, ctypes.byref(iterations), ctypes.byref(stop)):
        return None
    count = ctypes.c_uint64()
    recorded = recorded_at(result, ctypes.byref(count))
    size = int(count.value)
    steps = np.array(np.ctypeslib.as_array(recorded, shape=(size,)), dtype=np.int64) if size else np.zeros(0, np.int64)
    values = [_checked(values_of(result, which, ctypes.byref(count)), count, (size,)) for which in range(5)]
    seconds, primal, dual, scaling, partial_gap = values
    return Solution(int(iterations.value), Stop(int(stop.value)), History(steps, seconds, primal, dual, scaling,
                                                                          partial_gap))


def _times(library: ctypes.CDLL, result: ctypes.c_void_p) -> KernelTimes | None:
    values = (ctypes.c_double * 6)()
    if not library.unround_result_kernel_times(result, values, 6):
        return None
    return KernelTimes(values[0], int(values[1]), values[2], int(values[3]), values[4], int(values[5]))


def _decoded(library: ctypes.CDLL, result: ctypes.c_void_p) -> Decoded:
    height, width, channels = (ctypes.c_uint64() for _ in range(3))
    samples = library.unround_result_picture(result, ctypes.byref(height), ctypes.byref(width), ctypes.byref(channels))
    rows, columns, depth = int(height.value), int(width.value), int(channels.value)
    picture = _copy(samples, (rows, columns) if depth == 1 else (rows, columns, depth))
    planes_count, canvas_channels, canvas_height, canvas_width = (ctypes.c_uint64() for _ in range(4))
    canvas_pointer = library.unround_result_canvas(result, ctypes.byref(canvas_channels), ctypes.byref(canvas_height),
                                                   ctypes.byref(canvas_width))
    planes_pointer = library.unround_result_planes(result, ctypes.byref(planes_count))
    count = int(canvas_channels.value)
    planes = _checked(planes_pointer, planes_count, (count, rows, columns))
    canvas = _copy(canvas_pointer, (count, int(canvas_height.value), int(canvas_width.value)))
    components = [_problem(library, result, index) for index in range(int(library.unround_result_components(result)))]
    excess = []
    for index, (coefficients, _) in enumerate(components):
        count = ctypes.c_uint64()
        pointer = library.unround_result_excess(result, index, ctypes.byref(count))
        if not pointer:
            break
        excess.append(_checked(pointer, count, coefficients.shape))
    profile_size = ctypes.c_uint64()
    profile = library.unround_result_icc_profile(result, ctypes.byref(profile_size))
    warning = (library.unround_result_warning(result) or b"").decode("utf-8", "replace")
    if warning:
        warnings.warn(warning, UnroundWarning, stacklevel=4)
    return Decoded(
        picture=picture,
        planes=planes,
        canvas=canvas,
        color_space=jpegio.ColorSpace(int(library.unround_result_color_space(result))),
        coefficients=tuple(coefficients for coefficients, _ in components),
        problems=tuple(problem for _, problem in components),
        result=_solution(library, result),
        icc_profile=ctypes.string_at(profile, int(profile_size.value)) if profile else None,
        exif_orientation=int(library.unround_result_exif_orientation(result)),
        input_format="png" if int(library.unround_result_format(result)) == _FORMAT_PNG else "jpeg",
        bits=int(library.unround_result_bits(result)),
        alpha_result=_solution(library, result, alpha=True),
        excess=tuple(excess) if components and len(excess) == len(components) else None,
        kernel_times=_times(library, result),
    )


class _Watcher:
    """The observer as the interface calls it: each record made a Record, an exception kept."""

    def __init__(self, observer: Observer | None) -> None:
        self._observer = observer
        self.error: BaseException | None = None
        self.callback = _OBSERVER() if observer is None else _OBSERVER(self._call)

    def _call(self, _user: int | None, pointer: "ctypes._Pointer[_Record]") -> int:
        if self._observer is None or self.error is not None:
            return 1
        try:
            record = pointer.contents
            shape = (int(record.channels), int(record.height), int(record.width))
            seen = Record(int(record.iteration), float(record.gap), float(record.primal), float(record.dual),
                          _copy(record.canvas, shape))
            return 1 if self._observer(seen) else 0
        except BaseException as error:  # noqa: BLE001 -- raised again once the library has returned
            self.error = error
            return 1


def _reconstruct(call: Callable, observer: Observer | None) -> Decoded:
    library = _library()
    watcher = _Watcher(observer)
    result = ctypes.c_void_p()
    message = ctypes.create_string_buffer(_MESSAGE_SIZE)
    status = call(watcher.callback, ctypes.byref(result), message)
    try:
        if watcher.error is not None:
            raise watcher.error
        if status != Status.OK:
            known = status in Status.__members__.values()
            raise UnroundError(Status(status) if known else Status.INTERNAL, message.value.decode("utf-8", "replace"))
        return _decoded(library, result)
    finally:
        library.unround_result_free(result)


def decode(source: bytes | str | os.PathLike | jpegio.Image, settings: Settings | list[str] | None = None, *,
           observer: Observer | None = None, max_pixels: int = 0, max_scans: int = 0,
           warnings_are_errors: bool = False) -> Decoded:
    """Reconstructs a JPEG file or a PNG file, its bytes or its path (python/src/unround/
    native.py:677-716): a PNG file within the intervals of its rounded samples, by TV or TGV
    with mu and slack alone of the data term, its colours together and its alpha alone; any
    other as a JPEG file, read by the library's reader with these limits of reading. An Image
    is reconstructed as solve() does. A warning of the reading is said as an UnroundWarning."""
    if isinstance(source, jpegio.Image):
        if max_pixels or max_scans or warnings_are_errors:
            raise ValueError("an image read already takes no limits of reading")
        return solve(source, settings, observer=observer)
    library = _library()
    handle = _Settings(options(settings, max_pixels=max_pixels, max_scans=max_scans,
                               warnings_are_errors=warnings_are_errors))
    buffer = Path(source).read_bytes() if isinstance(source, (str, os.PathLike)) else bytes(source)

    def call(callback: object, result: object, message: ctypes.Array) -> int:
        return int(library.unround_decode(buffer, len(buffer), handle.handle, callback, None, result, message,
                                          _MESSAGE_SIZE))

    return _reconstruct(call, observer)


def project(picture: npt.ArrayLike, source: bytes | str | os.PathLike, settings: Settings | list[str] | None = None,
            *, max_pixels: int = 0, max_scans: int = 0, warnings_are_errors: bool = False) -> Decoded:
    """Measures a picture made elsewhere against a JPEG file's intervals and projects it onto
    them on the GPU (docs/math.md, 12; python/src/unround/native.py:718-755): picture (height,
    width) of a grey file or (height, width, 3) RGB of a file in colour, as the file stores
    it; the result is the projection as a reconstruction without a solver (result None), and
    its excess how far each coefficient lay outside its interval, in steps."""
    library = _library()
    samples, height, width, channels = _samples(picture)
    handle = _Settings(options(settings, max_pixels=max_pixels, max_scans=max_scans,
                               warnings_are_errors=warnings_are_errors))
    buffer = Path(source).read_bytes() if isinstance(source, (str, os.PathLike)) else bytes(source)
    pointer = samples.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def call(_callback: object, result: object, message: ctypes.Array) -> int:
        return int(library.unround_project(buffer, len(buffer), handle.handle, pointer, height, width, channels,
                                           result, message, _MESSAGE_SIZE))

    projected = _reconstruct(call, None)
    del samples
    return projected


def _input_of(image: jpegio.Image) -> tuple[_Input, list[np.ndarray], object]:
    count = len(image.components)
    levels = [np.ascontiguousarray(component.coefficients, dtype=np.int16) for component in image.components]
    raw = (_Component * max(count, 1))()
    for index, (component, values) in enumerate(zip(image.components, levels, strict=True)):
        if values.ndim != 4 or values.shape[2:] != (8, 8):
            raise ValueError(f"the levels of component {index} are (rows, columns, 8, 8), not {values.shape}")
        raw[index].levels = values.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))
        raw[index].rows, raw[index].columns = int(values.shape[0]), int(values.shape[1])
        raw[index].h_samp_factor = int(component.h_samp_factor)
        raw[index].v_samp_factor = int(component.v_samp_factor)
        table = np.asarray(component.quant_table, dtype=np.uint16).reshape(_BLOCK_SIZE)
        raw[index].quant_table = (ctypes.c_uint16 * _BLOCK_SIZE)(*(int(step) for step in table))
    return _Input(int(image.height), int(image.width), int(image.color_space), count, raw), levels, raw


def solve(image: jpegio.Image, settings: Settings | list[str] | None = None, *, observer: Observer | None = None,
          start: dict | None = None) -> Decoded:
    """Reconstructs the components of an image on the GPU (python/src/unround/native.py:758-788).

    settings may also be the command line's options themselves. start, this edition's,
    resumes from a state: 'coefficients' (one (rows, columns, 8, 8) array per component),
    'p' and 'w' (C, 2, H, W) and 'r' (C, 3, H, W), any of them left out.
    """
    library = _library()
    handle = _Settings(settings if isinstance(settings, list) else options(settings))
    given, levels, raw = _input_of(image)
    kept: list = [levels, raw]
    start_struct = None
    if start is not None:
        start_struct = _Start()
        coefficient_list = start.get("coefficients")
        if coefficient_list is not None:
            arrays = [np.ascontiguousarray(values, dtype=np.float64) for values in coefficient_list]
            pointers = (ctypes.POINTER(ctypes.c_double) * len(arrays))(
                *(array.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for array in arrays))
            kept += [arrays, pointers]
            start_struct.coefficients = ctypes.cast(pointers, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
        for name in ("p", "w", "r"):
            if start.get(name) is not None:
                array = np.ascontiguousarray(start[name], dtype=np.float64)
                kept.append(array)
                setattr(start_struct, name, array.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def call(callback: object, result: object, message: ctypes.Array) -> int:
        if start_struct is None:
            return int(library.unround_solve(ctypes.byref(given), handle.handle, callback, None, result, message,
                                             _MESSAGE_SIZE))
        return int(library.unround_solve_start(ctypes.byref(given), handle.handle, ctypes.byref(start_struct),
                                               callback, None, result, message, _MESSAGE_SIZE))

    decoded = _reconstruct(call, observer)
    del kept
    return decoded


def solve_samples(levels: npt.ArrayLike, bits: int = 8, settings: Settings | list[str] | None = None, *,
                  observer: Observer | None = None) -> Decoded:
    """Reconstructs a picture of rounded samples on the GPU, as decode() does a PNG file
    (rust/unround/src/rounded.rs:71-170): levels (height, width) or (height, width, channels)
    of `bits` bits, channels 1 grey, 2 grey and alpha, 3 RGB, 4 RGB and alpha; TV or TGV
    within the intervals of the levels with --mu and --slack alone of the data term, the
    colours together and the alpha alone. The Decoded is a PNG file's: input_format "png",
    the bits, no components, and alpha_result where there is alpha."""
    library = _library()
    handle = _Settings(settings if isinstance(settings, list) else options(settings))
    given = np.asarray(levels)
    if given.ndim not in (2, 3):
        raise ValueError(f"levels are (height, width) or (height, width, channels), not {given.shape}")
    if given.size and (np.min(given) < 0 or np.max(given) > 65535):
        raise ValueError("levels lie in 0..65535")
    values = np.ascontiguousarray(given, dtype=np.uint16)
    height, width = int(values.shape[0]), int(values.shape[1])
    channels = 1 if values.ndim == 2 else int(values.shape[2])

    def call(callback: object, result: object, message: ctypes.Array) -> int:
        return int(library.unround_solve_samples(values.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), height, width,
                                                 channels, bits, handle.handle, callback, None, result, message,
                                                 _MESSAGE_SIZE))

    return _reconstruct(call, observer)


def solve_batch(images: Sequence[jpegio.Image], settings: Settings | list[str] | None = None
                ) -> tuple[list[Decoded], list[float]]:
    """Reconstructs many images together on the GPU, this edition's: each result is what
    solve() gives for its image (the same iterates, records and stopping iteration), with
    the tiles of every image in the same launches and the stopping tests on the device. Also
    returns each image's latency: the seconds from the batch's first iteration to the end of
    the record at which it stopped (or of the last run of iterations). No observer."""
    library = _library()
    handle = _Settings(settings if isinstance(settings, list) else options(settings))
    count = len(images)
    inputs = (_Input * max(count, 1))()
    kept: list = []
    for index, image in enumerate(images):
        given, levels, raw = _input_of(image)
        inputs[index] = given
        kept += [levels, raw]
    results = (ctypes.c_void_p * max(count, 1))()
    latencies = np.zeros(max(count, 1))
    message = ctypes.create_string_buffer(_MESSAGE_SIZE)
    status = int(library.unround_solve_batch(inputs, count, handle.handle, results,
                                             latencies.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), message,
                                             _MESSAGE_SIZE))
    if status != Status.OK:
        known = status in Status.__members__.values()
        raise UnroundError(Status(status) if known else Status.INTERNAL, message.value.decode("utf-8", "replace"))
    decoded: list[Decoded] = []
    try:
        for index in range(count):
            decoded.append(_decoded(library, ctypes.c_void_p(results[index])))
    finally:
        for index in range(count):
            library.unround_result_free(ctypes.c_void_p(results[index]))
    del kept
    return decoded, [float(value) for value in latencies[:count]]


def _samples(picture: npt.ArrayLike) -> tuple[Array, int, int, int]:
    samples = np.ascontiguousarray(picture, dtype=np.float64)
    if samples.ndim == 2:
        return samples, int(samples.shape[0]), int(samples.shape[1]), 1
    if samples.ndim == 3:
        return samples, int(samples.shape[0]), int(samples.shape[1]), int(samples.shape[2])
    raise ValueError(f"a picture is (height, width) or (height, width, channels), not {samples.shape}")


def _written(name: str, picture: npt.ArrayLike, *arguments: object) -> bytes:
    """The file a writer of the library makes of the picture, with the arguments after its size;
    its warning, where it has one, as an UnroundWarning (python/src/unround/native.py:801-822)."""
    library = _library()
    samples, height, width, channels = _samples(picture)
    pointer = ctypes.POINTER(ctypes.c_uint8)()
    size = ctypes.c_uint64()
    message = ctypes.create_string_buffer(_MESSAGE_SIZE)
    status = getattr(library, name)(samples.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), height, width, channels,
                                    *arguments, ctypes.byref(pointer), ctypes.byref(size), message, _MESSAGE_SIZE)
    text = message.value.decode("utf-8", "replace")
    if status != Status.OK:
        known = status in Status.__members__.values()
        raise UnroundError(Status(status) if known else Status.INTERNAL, text)
    try:
        data = ctypes.string_at(pointer, int(size.value))
    finally:
        library.unround_bytes_free(pointer, size)
    if text:
        warnings.warn(text, UnroundWarning, stacklevel=3)
    return data


def _profile(icc_profile: bytes | None) -> tuple[bytes | None, int]:
    return (None, 0) if icc_profile is None else (bytes(icc_profile), len(icc_profile))


def tiff_bytes(picture: npt.ArrayLike, *, ycbcr: bool = False, icc_profile: bytes | None = None) -> bytes:
    """A TIFF of the picture's samples in binary64, bit for bit (python/src/unround/native.py:829-844):
    (height, width) or (height, width, channels), 1 to 4, alpha last. With ycbcr the three samples of
    a pixel are JFIF's Y, Cb and Cr, and the file says so. The ICC profile is embedded where it goes
    with the picture; otherwise it is left out, with an UnroundWarning."""
    return _written("unround_tiff", picture, 1 if ycbcr else 0, *_profile(icc_profile))


_DITHERS: Final = {("none", None): 0, ("blue-noise", None): 1, ("blue-noise", "apart"): 1,
                   ("blue-noise", "together"): 2}


def _dither(dither: str, channels: str | None) -> int:
    found = _DITHERS.get((dither, channels))
    if found is None:
        raise ValueError(f"the dither is blue-noise (channels apart or together) or none, not {dither}"
                         + ("" if channels is None else f" with the channels {channels}"))
    return found


def png_bytes(picture: npt.ArrayLike, *, sixteen: bool = False, dither: str = "blue-noise",
              dither_channels: str | None = None, compression: int | None = None,
              icc_profile: bytes | None = None) -> bytes:
    """A PNG file of the picture's samples, of 8 or 16 bits (python/src/unround/native.py:858-879):
    dither "blue-noise" (the ordered dither of the blue-noise matrix; dither_channels "apart", the
    default, or "together") or "none" (each sample rounded, halves away from 0) (docs/math.md, 10);
    compression zlib's level 0 to 9, None for zlib's default; the ICC profile as tiff_bytes takes it."""
    level = -1 if compression is None else compression
    if not -1 <= level <= 9:
        raise ValueError(f"zlib's level is 0 to 9, or None, not {compression}")
    return _written("unround_png", picture, 1 if sixteen else 0, _dither(dither, dither_channels), level,
                    *_profile(icc_profile))


def pnm_bytes(picture: npt.ArrayLike, *, sixteen: bool = False, dither: str = "blue-noise",
              dither_channels: str | None = None) -> bytes:
    """A PNM file (P5 grey, P6 RGB) of the picture's samples, made as png_bytes makes them."""
    return _written("unround_pnm", picture, 1 if sixteen else 0, _dither(dither, dither_channels))


def oriented(picture: npt.ArrayLike, orientation: int) -> Array:
    """The picture, (height, width) or (height, width, channels), turned upright as EXIF's
    orientation (1 to 8, 0 for none) says; from 5 on the rows become columns. Every sample moves
    bit for bit (python/src/unround/native.py:889-919)."""
    library = _library()
    samples, height, width, channels = _samples(picture)
    if not -(2**31) <= orientation < 2**31:
        raise ValueError(f"an EXIF orientation is 1 to 8, or 0 for none, not {orientation}")
    turned = np.empty(samples.size, dtype=np.float64)
    rows, columns = ctypes.c_uint64(), ctypes.c_uint64()
    message = ctypes.create_string_buffer(_MESSAGE_SIZE)
    status = library.unround_orient(samples.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), height, width, channels,
                                    orientation, turned.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    ctypes.byref(rows), ctypes.byref(columns), message, _MESSAGE_SIZE)
    if status != Status.OK:
        raise UnroundError(Status(status), message.value.decode("utf-8", "replace"))
    shape = (int(rows.value), int(columns.value)) + (() if samples.ndim == 2 else (channels,))
    return turned.reshape(shape)


def to_rgb(planes: npt.ArrayLike) -> Array:
    """JFIF's RGB (height, width, 3) of Y, Cb and Cr planes (3, height, width) (docs/math.md, 8)."""
    values = np.ascontiguousarray(planes, dtype=np.float64)
    _, height, width = values.shape
    out = np.zeros((height, width, 3))
    _library().unround_to_rgb(values.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), height, width,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def to_ycbcr(picture: npt.ArrayLike) -> Array:
    """JFIF's Y, Cb and Cr planes (3, height, width) of an RGB picture (height, width, 3)."""
    values = np.ascontiguousarray(picture, dtype=np.float64)
    height, width, _ = values.shape
    out = np.zeros((3, height, width))
    _library().unround_to_ycbcr(values.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), height, width,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out
