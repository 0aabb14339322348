"""JPEG-Unround on an AMD Instinct GPU, reached through its C interface (include/unround.h).

The public names, signatures and result types are the reference package's
(python/src/unround/native.py: options, held_options, decode, project, solve, the writers,
to_rgb and to_ycbcr, and the Decoded, Problem, History, Solution and Record types), so that
its callers carry over; this edition adds solve_batch, solve_samples and KernelTimes.

How this module is put together:

* the C layouts of the interface, and its entry points declared from one text table
  (_ENTRY_POINTS), each line a function's name, its result and its arguments in a small
  vocabulary of C types (_C_TYPES);
* _Call, one call of a function that reports a status and a message: it gives the message
  buffer and turns a failing status into UnroundError;
* Marshalled, an Image's components laid out as the interface's input, the arrays kept
  alive for as long as the object is;
* _ResultReader, one result handle read into arrays and the result types;
* the option writer, which turns Settings into the command line's options.

The library is the file that UNROUND_LIBRARY names, or else unround/_native/
libunround_capi.so, which jpeg-unround_amd/build.py writes. Without it every function
raises LibraryNotFoundError: nothing here computes on the CPU. ctypes lets go of the GIL
while the library runs.
"""

from __future__ import annotations

import ctypes
import enum
import os
import threading
import warnings
from collections.abc import Callable, Iterable
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
    "library_path",
    "options",
    "oriented",
    "png_bytes",
    "pnm_bytes",
    "project",
    "solve",
    "solve_batch",
    "solve_samples",
    "tiff_bytes",
    "to_rgb",
    "to_ycbcr",
    "version",
]

Array = npt.NDArray[np.float64]
Observer = Callable[["Record"], "bool | None"]

ABI_VERSION: Final = 5
MESSAGE_BYTES: Final = 4096
_PNG_FORMAT: Final = 2


# ---------------------------------------------------------------------------------------
# The interface's structures (include/unround.h; sizes checked when the library loads).

class CRecord(ctypes.Structure):
    """unround_record: what the observer is given at each record."""

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


class CComponent(ctypes.Structure):
    """unround_component: one component's levels, blocks, sampling and table."""

    _fields_ = (
        ("levels", ctypes.POINTER(ctypes.c_int16)),
        ("rows", ctypes.c_uint32),
        ("columns", ctypes.c_uint32),
        ("h_samp_factor", ctypes.c_int32),
        ("v_samp_factor", ctypes.c_int32),
        ("quant_table", ctypes.c_uint16 * 64),
    )


class CInput(ctypes.Structure):
    """unround_input: a picture's size, color space and components."""

    _fields_ = (
        ("height", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("color_space", ctypes.c_int32),
        ("count", ctypes.c_uint32),
        ("components", ctypes.POINTER(CComponent)),
    )


class CStart(ctypes.Structure):
    """unround_start (include/unround_ext.h): a state to resume from, null where not given."""

    _fields_ = (
        ("coefficients", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
        ("p", ctypes.POINTER(ctypes.c_double)),
        ("w", ctypes.POINTER(ctypes.c_double)),
        ("r", ctypes.POINTER(ctypes.c_double)),
    )


LAYOUT_BYTES: Final = {CRecord: 64, CComponent: 152, CInput: 24}

OBSERVER_TYPE: Final = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(CRecord))


# ---------------------------------------------------------------------------------------
# Errors and enumerations.

class Status(enum.IntEnum):
    """What a function of the interface reports (unround_status)."""

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
    """A call the library refused or could not finish: its status and its message."""

    def __init__(self, status: Status, message: str) -> None:
        super().__init__(f"{message} ({status.name.lower()})")
        self.status = status
        self.message = message

    def __reduce__(self) -> tuple[type["UnroundError"], tuple[Status, str]]:
        return (type(self), (self.status, self.message))


class LibraryNotFoundError(RuntimeError):
    """No library where this module looks for one, or one of another ABI version."""


class UnroundWarning(UserWarning):
    """A warning the library gave with a result."""


def status_error(status: int, text: str) -> UnroundError:
    """The error for a status other than OK; a status outside the enumeration is INTERNAL."""
    known = {member.value for member in Status}
    return UnroundError(Status(status) if status in known else Status.INTERNAL, text)


# ---------------------------------------------------------------------------------------
# The entry points.

_C_TYPES: Final[dict[str, object]] = {
    "void": None,
    "i32": ctypes.c_int32,
    "u32": ctypes.c_uint32,
    "u64": ctypes.c_uint64,
    "size": ctypes.c_size_t,
    "f64": ctypes.c_double,
    "text": ctypes.c_char_p,
    "texts": ctypes.POINTER(ctypes.c_char_p),
    "handle": ctypes.c_void_p,
    "handle*": ctypes.POINTER(ctypes.c_void_p),
    "address": ctypes.c_void_p,
    "f64*": ctypes.POINTER(ctypes.c_double),
    "u64*": ctypes.POINTER(ctypes.c_uint64),
    "i32*": ctypes.POINTER(ctypes.c_int32),
    "u8*": ctypes.POINTER(ctypes.c_uint8),
    "u8**": ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
    "u16*": ctypes.POINTER(ctypes.c_uint16),
    "input*": ctypes.POINTER(CInput),
    "start*": ctypes.POINTER(CStart),
    "observer": OBSERVER_TYPE,
}

# name, result, arguments; "message" stands for the pair (char* message, size_t size).
_ENTRY_POINTS: Final = """
unround_abi_version                     i32
unround_version                         text
unround_settings_new                    handle  size texts message
unround_settings_free                   void    handle
unround_settings_options                size    handle text size
unround_decode                          i32     text size handle observer handle handle* message
unround_solve                           i32     input* handle observer handle handle* message
unround_solve_start                     i32     input* handle start* observer handle handle* message
unround_solve_batch                     i32     input* size handle handle* f64* message
unround_solve_scratch                   i32     input* handle f64* u64 u64* i32* message
unround_solve_samples                   i32     u16* u64 u64 u64 u32 handle observer handle handle* message
unround_project                         i32     text size handle f64* u64 u64 u64 handle* message
unround_result_free                     void    handle
unround_result_picture                  f64*    handle u64* u64* u64*
unround_result_planes                   f64*    handle u64*
unround_result_canvas                   f64*    handle u64* u64* u64*
unround_result_color_space              i32     handle
unround_result_components               u64     handle
unround_result_component                i32     handle u64 u64* u64* u64* u64*
unround_result_coefficients             f64*    handle u64 u64*
unround_result_problem                  f64*    handle u64 i32 u64*
unround_result_excess                   f64*    handle u64 u64*
unround_result_solver                   i32     handle u64* i32*
unround_result_record_iterations        u64*    handle u64*
unround_result_record_values            f64*    handle i32 u64*
unround_result_alpha_solver             i32     handle u64* i32*
unround_result_alpha_record_iterations  u64*    handle u64*
unround_result_alpha_record_values      f64*    handle i32 u64*
unround_result_kernel_times             i32     handle f64* size
unround_result_exif_orientation         i32     handle
unround_result_icc_profile              u8*     handle u64*
unround_result_format                   i32     handle
unround_result_bits                     i32     handle
unround_result_warning                  text    handle
unround_tiff                            i32     f64* u64 u64 u64 i32 text u64 u8** u64* message
unround_png                             i32     f64* u64 u64 u64 i32 i32 i32 text u64 u8** u64* message
unround_pnm                             i32     f64* u64 u64 u64 i32 i32 u8** u64* message
unround_bytes_free                      void    u8* u64
unround_orient                          i32     f64* u64 u64 u64 i32 f64* u64* u64* message
unround_to_rgb                          i32     f64* u64 u64 f64*
unround_to_ycbcr                        i32     f64* u64 u64 f64*
"""


def declare(library: ctypes.CDLL, table: str) -> None:
    """Declares every function of a table written as _ENTRY_POINTS is."""
    for line in table.strip().splitlines():
        name, result, *arguments = line.split()
        words: list[str] = []
        for argument in arguments:
            words += ["text", "size"] if argument == "message" else [argument]
        function = getattr(library, name)
        function.restype = _C_TYPES[result]
        function.argtypes = tuple(_C_TYPES[word] for word in words)


_loaded: ctypes.CDLL | None = None
_loading_lock = threading.Lock()


def _lib() -> ctypes.CDLL:
    """The library, loaded and declared at the first call; checked against this module's
    layouts and ABI version."""
    global _loaded
    with _loading_lock:
        if _loaded is not None:
            return _loaded
        for layout, wanted in LAYOUT_BYTES.items():
            if ctypes.sizeof(layout) != wanted:
                raise RuntimeError(f"this Python's {layout.__name__} takes {ctypes.sizeof(layout)} bytes; "
                                   f"the C layout takes {wanted}")
        path = _loading.configured("UNROUND_LIBRARY") or _loading.bundled(_loading.LIBRARY)
        if path is None:
            raise LibraryNotFoundError("no libunround_capi.so in unround/_native and no UNROUND_LIBRARY: "
                                       "run python jpeg-unround_amd/build.py")
        if not path.is_file():
            raise LibraryNotFoundError(f"UNROUND_LIBRARY is {path}, which is not a file")
        library = ctypes.CDLL(str(path))
        declare(library, _ENTRY_POINTS)
        abi = int(library.unround_abi_version())
        if abi != ABI_VERSION:
            raise LibraryNotFoundError(f"{path} speaks ABI {abi}; this module speaks ABI {ABI_VERSION}")
        _loaded = library
        return library


def library_path() -> str:
    """The file of the library in use."""
    return str(_lib()._name)


def version() -> str:
    """The library's own description of itself."""
    return _lib().unround_version().decode("utf-8")


class _Call:
    """One call of a function that reports a status, with a message buffer of its own."""

    def __init__(self) -> None:
        self.buffer = ctypes.create_string_buffer(MESSAGE_BYTES)

    @property
    def message(self) -> tuple[ctypes.Array, int]:
        return self.buffer, MESSAGE_BYTES

    @property
    def text(self) -> str:
        return self.buffer.value.decode("utf-8", "replace")

    def check(self, status: int) -> None:
        if int(status) != Status.OK:
            raise status_error(int(status), self.text)


# ---------------------------------------------------------------------------------------
# Settings as options.

def _shortest(value: float) -> str:
    """The shortest decimal that reads back as the same double."""
    return repr(float(value))


def _maybe(value: float | None) -> str | None:
    return None if value is None else _shortest(value)


# Each option of G, and how a DataTerm gives its value. Several terms (one per component)
# write one value where they agree and a comma-separated list where they do not.
_DATA_TERM_OPTIONS: Final[tuple[tuple[str, Callable[[DataTerm], str]], ...]] = (
    ("--mu", lambda t: "rule" if t.mu is None else _shortest(t.mu)),
    ("--mu-scale", lambda t: _shortest(t.mu_scale)),
    ("--mu-power", lambda t: _shortest(t.mu_power)),
    ("--mu-chroma", lambda t: _shortest(t.mu_chroma)),
    ("--weight-power", lambda t: _shortest(t.power)),
    ("--dc-weight", lambda t: _shortest(t.dc_weight)),
    ("--centres", lambda t: t.centres),
    ("--slack", lambda t: _shortest(t.slack)),
    ("--slack-cost", lambda t: _shortest(t.slack_cost)),
    ("--flat-weight", lambda t: _shortest(t.flat_weight)),
    ("--flat-levels", lambda t: str(int(t.flat_levels))),
    ("--flat-dc", lambda t: "any" if t.flat_dc is None else str(int(t.flat_dc))),
)

# The primal-dual method's options: always written, or written where not None.
_PDHG_ALWAYS: Final[tuple[tuple[str, Callable[[PdhgOptions], str]], ...]] = (
    ("--relative-tolerance", lambda o: _shortest(o.relative_tolerance)),
    ("--partial-tolerance", lambda o: _shortest(o.partial_tolerance)),
    ("--step-product", lambda o: _shortest(o.step_product)),
    ("--record-every", lambda o: str(int(o.record_every))),
    ("--free-radius", lambda o: _shortest(o.free_radius)),
    ("--threads", lambda o: str(int(o.threads))),
    ("--precision", lambda o: o.precision),
)
_PDHG_GIVEN: Final[tuple[tuple[str, Callable[[PdhgOptions], str | None]], ...]] = (
    ("--iterations", lambda o: None if o.iterations is None else str(int(o.iterations))),
    ("--tolerance", lambda o: _maybe(o.tolerance)),
    ("--step-ratio", lambda o: _maybe(o.step_ratio)),
    ("--relaxation", lambda o: _maybe(o.relaxation)),
    ("--norm-squared", lambda o: _maybe(o.norm_squared)),
    ("--partial-radius", lambda o: _maybe(o.partial_radius)),
)
_SUBGRADIENT_OPTIONS: Final[tuple[tuple[str, Callable[[SubgradientOptions], str]], ...]] = (
    ("--subgradient-iterations", lambda o: str(int(o.iterations))),
    ("--subgradient-record-every", lambda o: str(int(o.record_every))),
    ("--subgradient-step", lambda o: _shortest(o.step)),
    ("--subgradient-decay", lambda o: _shortest(o.decay)),
)


def _pairs(items: Iterable[tuple[str, str | None]]) -> list[str]:
    out: list[str] = []
    for name, value in items:
        if value is not None:
            out += [name, value]
    return out


def _data_term_options(data: DataTerm | tuple[DataTerm, ...]) -> list[str]:
    terms = (data,) if isinstance(data, DataTerm) else tuple(data)
    written = []
    for name, value_of in _DATA_TERM_OPTIONS:
        values = [value_of(term) for term in terms]
        written.append((name, values[0] if all(v == values[0] for v in values) else ",".join(values)))
    return _pairs(written)


def _settings_options(settings: Settings) -> list[str]:
    model: TV | TGV = settings.tgv if settings.method == "tgv" else settings.tv
    out = _pairs((("--method", settings.method), ("--alpha", _shortest(settings.tv.alpha)),
                  ("--alpha1", _shortest(settings.tgv.alpha1)), ("--alpha0", _shortest(settings.tgv.alpha0))))
    if model.channel_weights is not None:
        out += ["--channel-weights", ",".join(_shortest(weight) for weight in model.channel_weights)]
    out += ["--channels", "coupled" if model.coupled else "apart"]
    out += _data_term_options(settings.data)
    solver = settings.pdhg
    out += _pairs((name, value_of(solver)) for name, value_of in _PDHG_ALWAYS)
    out += _pairs((name, value_of(solver)) for name, value_of in _PDHG_GIVEN)
    if not solver.scale_with_weight:
        out.append("--no-weight-scaling")
    out += _pairs((name, value_of(settings.subgradient)) for name, value_of in _SUBGRADIENT_OPTIONS)
    if not settings.subgradient.momentum:
        out.append("--no-momentum")
    return out


def options(settings: Settings | list[str] | None = None, *, max_pixels: int = 0, max_scans: int = 0,
            warnings_are_errors: bool = False) -> list[str]:
    """The command line's options that are these settings, then the limits of reading a
    file (0: the library's own). The channels' weights and coupling written are the
    method's model's: TGV's for TGV, TV's otherwise. A list of options is taken as it is."""
    if max_pixels < 0 or max_scans < 0:
        raise ValueError(f"limits of reading cannot be negative: max_pixels={max_pixels}, max_scans={max_scans}")
    if isinstance(settings, list):
        out = list(settings)
    else:
        out = _settings_options(Settings() if settings is None else settings)
    if max_pixels:
        out += ["--max-pixels", str(int(max_pixels))]
    if max_scans:
        out += ["--max-scans", str(int(max_scans))]
    if warnings_are_errors:
        out.append("--warnings-are-errors")
    return out


class SettingsHandle:
    """The library's settings made from options; freed with this object."""

    def __init__(self, given: Sequence[str]) -> None:
        self._library = _lib()
        words = [word.encode("utf-8") for word in given]
        array = (ctypes.c_char_p * max(1, len(words)))(*words)
        call = _Call()
        self.handle = self._library.unround_settings_new(len(words), array, *call.message)
        if not self.handle:
            raise UnroundError(Status.OPTIONS, call.text)

    @classmethod
    def of(cls, settings: Settings | list[str] | None) -> "SettingsHandle":
        return cls(options(settings))

    def held(self) -> list[str]:
        size = int(self._library.unround_settings_options(self.handle, None, 0))
        room = ctypes.create_string_buffer(size + 1)
        self._library.unround_settings_options(self.handle, room, size + 1)
        return room.value.decode("ascii").split(" ")

    def __del__(self) -> None:
        if getattr(self, "handle", None):
            self._library.unround_settings_free(self.handle)
            self.handle = None


def held_options(settings: Settings | list[str] | None = None) -> list[str]:
    """Every value the library holds for these settings, as --name=value options."""
    return SettingsHandle.of(settings).held()


# ---------------------------------------------------------------------------------------
# Result types.

@dataclass(frozen=True, slots=True, eq=False)
class Record:
    """A record as an observer receives it: the gap per sample, the primal and dual values,
    and a copy of the canvas (C, H, W)."""

    iteration: int
    gap: float
    primal: float
    dual: float
    canvas: Array


@dataclass(frozen=True, slots=True, eq=False)
class Problem:
    """One component's model: the canvas's samples per component sample (down, across), the
    intervals' ends and the centres (rows, columns, 8, 8), the steps and weights (8, 8)."""

    ratio: tuple[int, int]
    lower: Array
    upper: Array
    centres: Array
    steps: Array
    weights: Array


@dataclass(frozen=True, slots=True, eq=False)
class History:
    """The records: the iteration of each, the seconds to it, the primal and dual values,
    TGV's scaling theta and the partial gap (NaN where not taken)."""

    iterations: npt.NDArray[np.int64]
    seconds: Array
    primal: Array
    dual: Array
    scaling: Array
    partial_gap: Array

    @property
    def gap(self) -> Array:
        """Primal less dual at each record."""
        return np.asarray(self.primal - self.dual, dtype=np.float64)


@dataclass(frozen=True, slots=True, eq=False)
class Solution:
    """The solver's iterations, the reason it stopped, and its records."""

    iterations: int
    stop: Stop
    history: History

    @property
    def converged(self) -> bool:
        """True where a tolerance stopped the solver."""
        return self.stop is Stop.CONVERGED


@dataclass(frozen=True, slots=True)
class KernelTimes:
    """HIP-event durations of a solve's kernels. UNROUND_TIMING=1 during the solve times the
    runs of iterations between records; UNROUND_TIMING=2 also every launch;
    UNROUND_TIMING_WARMUP=W leaves the first W iterations untimed. dual_rows: the dual
    step's rows per wavefront."""

    primal_seconds: float
    primal_launches: int
    dual_seconds: float
    dual_launches: int
    iterations_seconds: float
    iterations: int
    dual_rows: int = 0


@dataclass(frozen=True, slots=True, eq=False)
class Decoded:
    """A reconstruction. picture: (height, width) or (height, width, C), binary64, JFIF's
    RGB of a YCbCr file; planes: (C, height, width); canvas: the solver's whole (C, H, W);
    coefficients and problems per component (none of a PNG file); result: the solver's
    (None for the decoder of the centres or a projection); alpha_result: a PNG file's
    alpha's; excess: a projection's distances outside the intervals, in steps."""

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


def _floats(pointer: object, shape: tuple[int, ...], what: str, count: int | None = None) -> Array:
    """A copy of the doubles at a pointer of the library as an array of `shape`; `count` is
    how many the library says it holds."""
    size = int(np.prod(shape, dtype=np.int64))
    if count is not None and int(count) != size:
        raise RuntimeError(f"{what}: the library holds {count} values, and shape {shape} needs {size}")
    if size == 0:
        return np.zeros(shape, dtype=np.float64)
    if not pointer:
        raise RuntimeError(f"{what}: the library returned no pointer for {size} values")
    return np.ctypeslib.as_array(pointer, shape=(size,)).astype(np.float64, copy=True).reshape(shape)


class _HeldResult:
    """A result handle of the library, freed once nothing holds it: the arrays read from it
    view its memory rather than copy it, and each keeps this alive."""

    def __init__(self, library: ctypes.CDLL, handle: ctypes.c_void_p) -> None:
        self.lib = library
        self.handle = handle

    def free(self) -> None:
        handle, self.handle = self.handle, None
        if handle:
            self.lib.unround_result_free(handle)

    def __del__(self) -> None:
        try:
            self.free()
        except Exception:  # noqa: BLE001 -- the library may be gone at interpreter exit
            pass


def _viewed(pointer: object, shape: tuple[int, ...], what: str, count: int | None, held: _HeldResult) -> Array:
    """The doubles at a pointer into a held result as an array of `shape` over that memory
    (writable, no copy), which keeps the result alive; `count` is how many the library says
    it holds."""
    size = int(np.prod(shape, dtype=np.int64))
    if count is not None and int(count) != size:
        raise RuntimeError(f"{what}: the library holds {count} values, and shape {shape} needs {size}")
    if size == 0:
        return np.zeros(shape, dtype=np.float64)
    if not pointer:
        raise RuntimeError(f"{what}: the library returned no pointer for {size} values")
    memory = (ctypes.c_double * size).from_address(ctypes.cast(pointer, ctypes.c_void_p).value)
    memory._held = held  # the array's buffer keeps the result, which frees the handle last
    return np.frombuffer(memory, dtype=np.float64).reshape(shape)


class _ResultReader:
    """One result handle of the library, read into this module's types: the large arrays
    (picture, canvas, planes, coefficients, the problems, the records) view the result's
    memory, which lives until the last of them goes."""

    def __init__(self, library: ctypes.CDLL, handle: ctypes.c_void_p) -> None:
        self.lib = library
        self.handle = handle
        self.held = _HeldResult(library, handle)

    def _floats(self, pointer: object, shape: tuple[int, ...], what: str, count: int | None = None) -> Array:
        return _viewed(pointer, shape, what, count, self.held)

    def _counted(self, function: Callable, *arguments: object) -> tuple[object, int]:
        count = ctypes.c_uint64()
        pointer = function(self.handle, *arguments, ctypes.byref(count))
        return pointer, int(count.value)

    def picture(self) -> Array:
        sizes = [ctypes.c_uint64() for _ in range(3)]
        pointer = self.lib.unround_result_picture(self.handle, *map(ctypes.byref, sizes))
        height, width, channels = (int(size.value) for size in sizes)
        shape = (height, width) if channels == 1 else (height, width, channels)
        return self._floats(pointer, shape, "the picture")

    def canvas_and_planes(self, height: int, width: int) -> tuple[Array, Array]:
        sizes = [ctypes.c_uint64() for _ in range(3)]
        pointer = self.lib.unround_result_canvas(self.handle, *map(ctypes.byref, sizes))
        channels, rows, columns = (int(size.value) for size in sizes)
        canvas = self._floats(pointer, (channels, rows, columns), "the canvas")
        planes_pointer, planes_count = self._counted(self.lib.unround_result_planes)
        planes = self._floats(planes_pointer, (channels, height, width), "the planes", planes_count)
        return canvas, planes

    def component(self, index: int) -> tuple[Array, Problem]:
        sizes = [ctypes.c_uint64() for _ in range(4)]
        if self.lib.unround_result_component(self.handle, index, *map(ctypes.byref, sizes)) != Status.OK:
            raise RuntimeError(f"the library has no component {index} in this result")
        rows, columns, down, across = (int(size.value) for size in sizes)
        blocks = (rows, columns, 8, 8)
        pointer, count = self._counted(self.lib.unround_result_coefficients, index)
        coefficients = self._floats(pointer, blocks, f"coefficients of component {index}", count)
        arrays = []
        for which, (what, shape) in enumerate((("lower ends", blocks), ("upper ends", blocks), ("centres", blocks),
                                               ("steps", (8, 8)), ("weights", (8, 8)))):
            pointer, count = self._counted(self.lib.unround_result_problem, index, which)
            arrays.append(self._floats(pointer, shape, f"{what} of component {index}", count))
        return coefficients, Problem((down, across), *arrays)

    def solution(self, alpha: bool = False) -> Solution | None:
        prefix = "unround_result_alpha_" if alpha else "unround_result_"
        iterations, stop = ctypes.c_uint64(), ctypes.c_int32()
        if not getattr(self.lib, prefix + "solver")(self.handle, ctypes.byref(iterations), ctypes.byref(stop)):
            return None
        pointer, count = self._counted(getattr(self.lib, prefix + "record_iterations"))
        recorded = (np.ctypeslib.as_array(pointer, shape=(count,)).astype(np.int64, copy=True) if count
                    else np.zeros(0, dtype=np.int64))
        columns = []
        for which, what in enumerate(("seconds", "primal values", "dual values", "scalings", "partial gaps")):
            values, values_count = self._counted(getattr(self.lib, prefix + "record_values"), which)
            columns.append(self._floats(values, (count,), f"the records' {what}", values_count))
        return Solution(int(iterations.value), Stop(int(stop.value)), History(recorded, *columns))

    def excess(self, shapes: list[tuple[int, ...]]) -> tuple[Array, ...] | None:
        found = []
        for index, shape in enumerate(shapes):
            pointer, count = self._counted(self.lib.unround_result_excess, index)
            if not pointer:
                return None
            found.append(self._floats(pointer, shape, f"excess of component {index}", count))
        return tuple(found) if found else None

    def kernel_times(self) -> KernelTimes | None:
        values = (ctypes.c_double * 7)()
        if not self.lib.unround_result_kernel_times(self.handle, values, 7):
            return None
        return KernelTimes(values[0], int(values[1]), values[2], int(values[3]), values[4], int(values[5]),
                           int(values[6]))

    def icc_profile(self) -> bytes | None:
        pointer, count = self._counted(self.lib.unround_result_icc_profile)
        return ctypes.string_at(pointer, count) if pointer else None

    def decoded(self) -> Decoded:
        picture = self.picture()
        canvas, planes = self.canvas_and_planes(picture.shape[0], picture.shape[1])
        parts = [self.component(index) for index in range(int(self.lib.unround_result_components(self.handle)))]
        said = self.lib.unround_result_warning(self.handle)
        if said:
            warnings.warn(said.decode("utf-8", "replace"), UnroundWarning, stacklevel=4)
        return Decoded(
            picture=picture,
            planes=planes,
            canvas=canvas,
            color_space=jpegio.ColorSpace(int(self.lib.unround_result_color_space(self.handle))),
            coefficients=tuple(coefficients for coefficients, _ in parts),
            problems=tuple(problem for _, problem in parts),
            result=self.solution(),
            icc_profile=self.icc_profile(),
            exif_orientation=int(self.lib.unround_result_exif_orientation(self.handle)),
            input_format="png" if int(self.lib.unround_result_format(self.handle)) == _PNG_FORMAT else "jpeg",
            bits=int(self.lib.unround_result_bits(self.handle)),
            alpha_result=self.solution(alpha=True),
            excess=self.excess([coefficients.shape for coefficients, _ in parts]),
            kernel_times=self.kernel_times(),
        )


def read_result(handle: ctypes.c_void_p) -> Decoded:
    """The Decoded of a result handle, which is freed when the last of its arrays goes (at
    once where reading it fails)."""
    reader = _ResultReader(_lib(), handle)
    try:
        return reader.decoded()
    except BaseException:
        reader.held.free()
        raise


# ---------------------------------------------------------------------------------------
# Calls that make one result.

class _ObserverRelay:
    """The observer as the library calls it. Each record becomes a Record; an exception the
    observer raises stops the solver and is raised again once the library returns."""

    def __init__(self, observer: Observer | None) -> None:
        self.observer = observer
        self.raised: BaseException | None = None
        self.function = OBSERVER_TYPE(self) if observer is not None else OBSERVER_TYPE()

    def __call__(self, _user: object, pointer: "ctypes._Pointer[CRecord]") -> int:
        if self.raised is not None or self.observer is None:
            return 1
        try:
            c = pointer.contents
            record = Record(int(c.iteration), float(c.gap), float(c.primal), float(c.dual),
                            _floats(c.canvas, (int(c.channels), int(c.height), int(c.width)), "the record's canvas"))
            return 1 if self.observer(record) else 0
        except BaseException as error:  # noqa: BLE001 -- raised again after the library returns
            self.raised = error
            return 1


def _one_result(run: Callable[[object, ctypes.c_void_p, _Call], int], observer: Observer | None) -> Decoded:
    """Runs a call that makes one result (observer, result out, message) and reads it."""
    relay = _ObserverRelay(observer)
    handle = ctypes.c_void_p()
    call = _Call()
    status = int(run(relay.function, handle, call))
    if relay.raised is not None:
        if handle:
            _lib().unround_result_free(handle)
        raise relay.raised
    if status != Status.OK:
        if handle:
            _lib().unround_result_free(handle)
        raise status_error(status, call.text)
    return read_result(handle)


class Marshalled:
    """An Image's components as the interface's unround_input; the level arrays and the
    component table live as long as this object."""

    def __init__(self, image: jpegio.Image) -> None:
        self.levels = []
        self.table = (CComponent * max(1, len(image.components)))()
        for index, component in enumerate(image.components):
            levels = np.ascontiguousarray(component.coefficients, dtype=np.int16)
            if levels.ndim != 4 or levels.shape[2:] != (8, 8):
                raise ValueError(f"component {index}: levels must be (rows, columns, 8, 8); got {levels.shape}")
            self.levels.append(levels)
            entry = self.table[index]
            entry.levels = levels.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))
            entry.rows, entry.columns = levels.shape[0], levels.shape[1]
            entry.h_samp_factor, entry.v_samp_factor = int(component.h_samp_factor), int(component.v_samp_factor)
            steps = np.asarray(component.quant_table, dtype=np.uint16).ravel()
            entry.quant_table[:] = [int(step) for step in steps]
        self.input = CInput(int(image.height), int(image.width), int(image.color_space), len(image.components),
                            self.table)


class _StartState:
    """The state solve() resumes from, as unround_start, its arrays kept alive."""

    def __init__(self, start: dict) -> None:
        self.struct = CStart()
        self.kept: list[object] = []
        doubles = ctypes.POINTER(ctypes.c_double)
        coefficients = start.get("coefficients")
        if coefficients is not None:
            arrays = [np.ascontiguousarray(values, dtype=np.float64) for values in coefficients]
            table = (doubles * len(arrays))(*(array.ctypes.data_as(doubles) for array in arrays))
            self.kept += [arrays, table]
            self.struct.coefficients = ctypes.cast(table, ctypes.POINTER(doubles))
        for name in ("p", "w", "r"):
            if start.get(name) is not None:
                array = np.ascontiguousarray(start[name], dtype=np.float64)
                self.kept.append(array)
                setattr(self.struct, name, array.ctypes.data_as(doubles))


def _file_bytes(source: bytes | str | os.PathLike) -> bytes:
    return Path(source).read_bytes() if isinstance(source, (str, os.PathLike)) else bytes(source)


def decode(source: bytes | str | os.PathLike | jpegio.Image, settings: Settings | list[str] | None = None, *,
           observer: Observer | None = None, max_pixels: int = 0, max_scans: int = 0,
           warnings_are_errors: bool = False) -> Decoded:
    """Reconstructs a file (its bytes or its path) on the GPU: a PNG file within the
    intervals of its rounded samples (TV or TGV, mu and slack alone of the data term, the
    colours together and the alpha alone); any other file as a JPEG file, read with these
    limits. An Image is solved as solve() solves it. A warning of the reader is an
    UnroundWarning."""
    if isinstance(source, jpegio.Image):
        if max_pixels or max_scans or warnings_are_errors:
            raise ValueError("limits of reading apply to files; this Image has been read already")
        return solve(source, settings, observer=observer)
    library = _lib()
    held = SettingsHandle(options(settings, max_pixels=max_pixels, max_scans=max_scans,
                                  warnings_are_errors=warnings_are_errors))
    data = _file_bytes(source)
    return _one_result(lambda relay, handle, call: library.unround_decode(
        data, len(data), held.handle, relay, None, ctypes.byref(handle), *call.message), observer)


def _picture_array(picture: npt.ArrayLike) -> tuple[Array, int, int, int]:
    samples = np.ascontiguousarray(picture, dtype=np.float64)
    if samples.ndim not in (2, 3):
        raise ValueError(f"a picture has 2 or 3 axes (height, width[, channels]); got shape {samples.shape}")
    channels = 1 if samples.ndim == 2 else int(samples.shape[2])
    return samples, int(samples.shape[0]), int(samples.shape[1]), channels


def _doubles(array: np.ndarray) -> object:
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def project(picture: npt.ArrayLike, source: bytes | str | os.PathLike, settings: Settings | list[str] | None = None,
            *, max_pixels: int = 0, max_scans: int = 0, warnings_are_errors: bool = False) -> Decoded:
    """A picture made elsewhere, (height, width) grey or (height, width, 3) RGB, measured
    against a JPEG file's intervals and projected onto them on the GPU (docs/math.md 12):
    a Decoded without a solver whose excess holds how far each coefficient lay outside its
    interval, in steps."""
    library = _lib()
    samples, height, width, channels = _picture_array(picture)
    held = SettingsHandle(options(settings, max_pixels=max_pixels, max_scans=max_scans,
                                  warnings_are_errors=warnings_are_errors))
    data = _file_bytes(source)
    return _one_result(lambda _relay, handle, call: library.unround_project(
        data, len(data), held.handle, _doubles(samples), height, width, channels, ctypes.byref(handle),
        *call.message), None)


def solve(image: jpegio.Image, settings: Settings | list[str] | None = None, *, observer: Observer | None = None,
          start: dict | None = None) -> Decoded:
    """Reconstructs an Image's components on the GPU. settings may be the command line's
    options themselves. start (this edition's) resumes from a state: 'coefficients', one
    (rows, columns, 8, 8) array per component, 'p' and 'w' (C, 2, H, W), 'r' (C, 3, H, W);
    any of them may be left out."""
    library = _lib()
    held = SettingsHandle.of(settings)
    marshalled = Marshalled(image)
    state = None if start is None else _StartState(start)

    def run(relay: object, handle: ctypes.c_void_p, call: _Call) -> int:
        if state is None:
            return library.unround_solve(ctypes.byref(marshalled.input), held.handle, relay, None,
                                         ctypes.byref(handle), *call.message)
        return library.unround_solve_start(ctypes.byref(marshalled.input), held.handle, ctypes.byref(state.struct),
                                           relay, None, ctypes.byref(handle), *call.message)

    return _one_result(run, observer)


def solve_samples(levels: npt.ArrayLike, bits: int = 8, settings: Settings | list[str] | None = None, *,
                  observer: Observer | None = None) -> Decoded:
    """Reconstructs a picture of rounded samples on the GPU as decode() does a PNG file
    (rust/unround/src/rounded.rs:71-170): levels (height, width) or (height, width,
    channels) of `bits` bits, channels 1 grey, 2 grey and alpha, 3 RGB, 4 RGB and alpha.
    The Decoded is a PNG file's: no components, alpha_result where there is alpha."""
    library = _lib()
    held = SettingsHandle.of(settings)
    given = np.asarray(levels)
    if given.ndim not in (2, 3):
        raise ValueError(f"levels have 2 or 3 axes (height, width[, channels]); got shape {given.shape}")
    if given.size and not (0 <= np.min(given) and np.max(given) <= 0xFFFF):
        raise ValueError("every level must fit 16 bits (0 to 65535)")
    values = np.ascontiguousarray(given, dtype=np.uint16)
    height, width = int(values.shape[0]), int(values.shape[1])
    channels = 1 if values.ndim == 2 else int(values.shape[2])
    pointer = values.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
    return _one_result(lambda relay, handle, call: library.unround_solve_samples(
        pointer, height, width, channels, bits, held.handle, relay, None, ctypes.byref(handle), *call.message),
        observer)


def solve_batch(images: Sequence[jpegio.Image], settings: Settings | list[str] | None = None
                ) -> tuple[list[Decoded], list[float]]:
    """Reconstructs many images together on the GPU (this edition's): each Decoded is what
    solve() gives for its image, with every image's tiles in shared launches and the stopping
    tests on the device. Also gives each image's latency: the seconds from the batch's first
    iteration to the end of the run of iterations in which it stopped. No observer."""
    library = _lib()
    held = SettingsHandle.of(settings)
    marshalled = [Marshalled(image) for image in images]
    count = len(marshalled)
    inputs = (CInput * max(1, count))(*(m.input for m in marshalled))
    handles = (ctypes.c_void_p * max(1, count))()
    latencies = np.zeros(max(1, count), dtype=np.float64)
    call = _Call()
    call.check(library.unround_solve_batch(inputs, count, held.handle, handles, _doubles(latencies), *call.message))
    readers = [_ResultReader(library, ctypes.c_void_p(handles[index])) for index in range(count)]
    try:
        results = [reader.decoded() for reader in readers]
    except BaseException:
        for reader in readers:
            reader.held.free()
        raise
    return results, [float(seconds) for seconds in latencies[:count]]


# ---------------------------------------------------------------------------------------
# Writers and conversions.

def _file_of(function: str, picture: npt.ArrayLike, *arguments: object) -> bytes:
    """The bytes a writer of the library makes of a picture (its arguments after the
    picture's size); its warning, if any, as an UnroundWarning."""
    library = _lib()
    samples, height, width, channels = _picture_array(picture)
    made = ctypes.POINTER(ctypes.c_uint8)()
    size = ctypes.c_uint64()
    call = _Call()
    status = getattr(library, function)(_doubles(samples), height, width, channels, *arguments, ctypes.byref(made),
                                        ctypes.byref(size), *call.message)
    call.check(status)
    try:
        data = ctypes.string_at(made, int(size.value))
    finally:
        library.unround_bytes_free(made, size)
    if call.text:
        warnings.warn(call.text, UnroundWarning, stacklevel=3)
    return data


def _icc(profile: bytes | None) -> tuple[bytes | None, int]:
    return (None, 0) if profile is None else (bytes(profile), len(profile))


def tiff_bytes(picture: npt.ArrayLike, *, ycbcr: bool = False, icc_profile: bytes | None = None) -> bytes:
    """A TIFF holding the picture's samples as binary64, bit for bit: (height, width) or
    (height, width, channels), 1 to 4 channels, alpha last. With ycbcr the three samples of
    a pixel are JFIF's Y, Cb and Cr. An ICC profile that does not suit the picture is left
    out, with an UnroundWarning."""
    return _file_of("unround_tiff", picture, int(bool(ycbcr)), *_icc(icc_profile))


_DITHER_CODES: Final = {"none": {None: 0}, "blue-noise": {None: 1, "apart": 1, "together": 2}}


def _dither_code(dither: str, channels: str | None) -> int:
    code = _DITHER_CODES.get(dither, {}).get(channels)
    if code is None:
        raise ValueError(f"unknown dither {dither!r} with channels {channels!r}: use 'none', or 'blue-noise' "
                         "with channels 'apart' (the default) or 'together'")
    return code


def png_bytes(picture: npt.ArrayLike, *, sixteen: bool = False, dither: str = "blue-noise",
              dither_channels: str | None = None, compression: int | None = None,
              icc_profile: bytes | None = None) -> bytes:
    """A PNG file of the picture, 8 or 16 bits a sample (docs/math.md 10): dither
    "blue-noise" (ordered by the blue-noise ranks; dither_channels "apart", the default, or
    "together") or "none" (each sample rounded, halves away from 0); compression zlib's
    level 0 to 9, None for its default; the ICC profile as tiff_bytes takes it."""
    level = -1 if compression is None else int(compression)
    if level < -1 or level > 9:
        raise ValueError(f"compression is a zlib level from 0 to 9 (or None); got {compression}")
    return _file_of("unround_png", picture, int(bool(sixteen)), _dither_code(dither, dither_channels), level,
                    *_icc(icc_profile))


def pnm_bytes(picture: npt.ArrayLike, *, sixteen: bool = False, dither: str = "blue-noise",
              dither_channels: str | None = None) -> bytes:
    """A PNM file (P5 grey, P6 RGB) of the picture, its samples made as png_bytes makes them."""
    return _file_of("unround_pnm", picture, int(bool(sixteen)), _dither_code(dither, dither_channels))


def oriented(picture: npt.ArrayLike, orientation: int) -> Array:
    """The picture turned upright as an EXIF orientation (1 to 8; 0 for none) says; from 5
    on, rows and columns trade places. Samples move bit for bit."""
    library = _lib()
    samples, height, width, channels = _picture_array(picture)
    if not -(2**31) <= int(orientation) < 2**31:
        raise ValueError(f"an EXIF orientation is 0 to 8; got {orientation}")
    turned = np.empty(samples.size, dtype=np.float64)
    rows, columns = ctypes.c_uint64(), ctypes.c_uint64()
    call = _Call()
    call.check(library.unround_orient(_doubles(samples), height, width, channels, int(orientation), _doubles(turned),
                                      ctypes.byref(rows), ctypes.byref(columns), *call.message))
    shape = (int(rows.value), int(columns.value)) + (() if samples.ndim == 2 else (channels,))
    return turned.reshape(shape)


def _convert(function: str, given: npt.ArrayLike, axes: str) -> Array:
    values = np.ascontiguousarray(given, dtype=np.float64)
    if values.ndim != 3 or (values.shape[0] if axes == "planes" else values.shape[2]) != 3:
        wanted = "(3, height, width)" if axes == "planes" else "(height, width, 3)"
        raise ValueError(f"expected {axes} of shape {wanted}; got {values.shape}")
    height, width = (values.shape[1], values.shape[2]) if axes == "planes" else values.shape[:2]
    out = np.empty((height, width, 3) if axes == "planes" else (3, height, width))
    status = getattr(_lib(), function)(_doubles(values), height, width, _doubles(out))
    if status != Status.OK:
        raise status_error(int(status), f"{function} refused a picture of {height} x {width}")
    return out


def to_rgb(planes: npt.ArrayLike) -> Array:
    """JFIF's RGB, (height, width, 3), of Y, Cb and Cr planes (3, height, width) (docs/math.md 8)."""
    return _convert("unround_to_rgb", planes, "planes")


def to_ycbcr(picture: npt.ArrayLike) -> Array:
    """JFIF's Y, Cb and Cr planes, (3, height, width), of an RGB picture (height, width, 3)."""
    return _convert("unround_to_ycbcr", picture, "picture")
