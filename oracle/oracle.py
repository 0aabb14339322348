"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle/unround_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline import this module,
as the checker. It also reads option strings as the reference's command line names them
(cpp/src/settings.cpp:36-72, 242-410) into the oracle's structures; ranges that the
oracle itself does not check are checked by the library under test, not here.
"""

from __future__ import annotations

import ctypes
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "_build" / "liboracle.so"


class OracleComponent(ctypes.Structure):
    _fields_ = (
        ("levels", ctypes.POINTER(ctypes.c_int16)),
        ("rows", ctypes.c_uint32),
        ("columns", ctypes.c_uint32),
        ("h", ctypes.c_int32),
        ("v", ctypes.c_int32),
        ("table", ctypes.c_uint16 * 64),
    )


class OracleData(ctypes.Structure):
    _fields_ = (
        ("has_mu", ctypes.c_int32),
        ("mu", ctypes.c_double),
        ("mu_scale", ctypes.c_double),
        ("mu_power", ctypes.c_double),
        ("mu_chroma", ctypes.c_double),
        ("slack", ctypes.c_double),
        ("slack_cost", ctypes.c_double),
        ("dc_weight", ctypes.c_double),
        ("power", ctypes.c_double),
        ("flat_weight", ctypes.c_double),
        ("flat_levels", ctypes.c_uint32),
        ("has_flat_dc", ctypes.c_int32),
        ("flat_dc", ctypes.c_uint32),
        ("centres", ctypes.c_int32),
    )


class OracleOptions(ctypes.Structure):
    _fields_ = (
        ("method", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("alpha1", ctypes.c_double),
        ("alpha0", ctypes.c_double),
        ("coupled", ctypes.c_int32),
        ("has_channel_weights", ctypes.c_int32),
        ("channel_weights", ctypes.c_double * 3),
        ("has_iterations", ctypes.c_int32),
        ("iterations", ctypes.c_uint64),
        ("has_tolerance", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("relative_tolerance", ctypes.c_double),
        ("partial_tolerance", ctypes.c_double),
        ("has_partial_radius", ctypes.c_int32),
        ("partial_radius", ctypes.c_double),
        ("has_step_ratio", ctypes.c_int32),
        ("step_ratio", ctypes.c_double),
        ("has_relaxation", ctypes.c_int32),
        ("relaxation", ctypes.c_double),
        ("step_product", ctypes.c_double),
        ("has_norm_squared", ctypes.c_int32),
        ("norm_squared", ctypes.c_double),
        ("weight_scaling", ctypes.c_int32),
        ("record_every", ctypes.c_uint64),
        ("free_radius", ctypes.c_double),
    )


class OracleStart(ctypes.Structure):
    _fields_ = (
        ("coefficients", ctypes.POINTER(ctypes.c_double) * 3),
        ("p", ctypes.POINTER(ctypes.c_double)),
        ("w", ctypes.POINTER(ctypes.c_double)),
        ("r", ctypes.POINTER(ctypes.c_double)),
    )


def build() -> Path:
    """Compiles the oracle with its Makefile where its library is missing or older than its source."""
    source = HERE / "unround_oracle.c"
    if not LIBRARY.exists() or LIBRARY.stat().st_mtime < source.stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIBRARY


_library: ctypes.CDLL | None = None


def library() -> ctypes.CDLL:
    global _library
    if _library is None:
        if not LIBRARY.exists():
            build()
        lib = ctypes.CDLL(str(LIBRARY))
        dp, u64p, i32p = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32)
        lib.oracle_solve.restype = ctypes.c_void_p
        lib.oracle_solve.argtypes = (
            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
            ctypes.POINTER(OracleComponent), ctypes.POINTER(OracleData), ctypes.POINTER(OracleOptions),
            ctypes.POINTER(OracleStart), ctypes.c_char_p, ctypes.c_size_t,
        )
        lib.oracle_result_free.argtypes = (ctypes.c_void_p,)
        lib.oracle_result_canvas.restype = dp
        lib.oracle_result_canvas.argtypes = (ctypes.c_void_p, u64p, u64p, u64p)
        lib.oracle_result_iterate.restype = dp
        lib.oracle_result_iterate.argtypes = (ctypes.c_void_p,)
        lib.oracle_result_component.argtypes = (ctypes.c_void_p, ctypes.c_uint32, u64p, u64p, u64p, u64p)
        lib.oracle_result_array.restype = dp
        lib.oracle_result_array.argtypes = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, u64p)
        lib.oracle_result_solver.argtypes = (ctypes.c_void_p, dp, dp, u64p, i32p)
        lib.oracle_result_iteration_seconds.restype = ctypes.c_double
        lib.oracle_result_iteration_seconds.argtypes = (ctypes.c_void_p,)
        lib.oracle_result_least_iteration_seconds.restype = ctypes.c_double
        lib.oracle_result_least_iteration_seconds.argtypes = (ctypes.c_void_p,)
        lib.oracle_set_threads.argtypes = (ctypes.c_int32,)
        lib.oracle_threads.restype = ctypes.c_int32
        lib.oracle_result_record_iterations.restype = u64p
        lib.oracle_result_record_iterations.argtypes = (ctypes.c_void_p, u64p)
        lib.oracle_result_record_values.restype = dp
        lib.oracle_result_record_values.argtypes = (ctypes.c_void_p, ctypes.c_int32, u64p)
        lib.oracle_basis.argtypes = (dp,)
        lib.oracle_dct_forward.argtypes = (dp, dp)
        lib.oracle_dct_inverse.argtypes = (dp, dp)
        lib.oracle_laplace_scale.restype = ctypes.c_double
        lib.oracle_laplace_scale.argtypes = (ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double)
        lib.oracle_laplace_shrinkage.restype = ctypes.c_double
        lib.oracle_laplace_shrinkage.argtypes = (ctypes.c_double,)
        lib.oracle_mmse_centre.restype = ctypes.c_double
        lib.oracle_mmse_centre.argtypes = (ctypes.c_int16, ctypes.c_double, ctypes.c_double)
        lib.oracle_sum_rows.restype = ctypes.c_double
        lib.oracle_sum_rows.argtypes = (dp, ctypes.c_uint64, ctypes.c_uint64)
        u16p = ctypes.POINTER(ctypes.c_uint16)
        lib.oracle_samples_intervals.restype = ctypes.c_int
        lib.oracle_samples_intervals.argtypes = (u16p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, dp, dp, dp)
        lib.oracle_samples_prox.argtypes = (u16p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                            dp, ctypes.c_double, dp)
        for name in ("oracle_samples_value", "oracle_samples_conjugate"):
            getattr(lib, name).restype = ctypes.c_double
            getattr(lib, name).argtypes = (u16p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_double, ctypes.c_double, dp)
        lib.oracle_samples_solve.restype = ctypes.c_void_p
        lib.oracle_samples_solve.argtypes = (u16p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_double, ctypes.c_double, ctypes.POINTER(OracleOptions),
                                             ctypes.c_char_p, ctypes.c_size_t)
        _library = lib
    return _library


def _doubles(array: np.ndarray) -> ctypes.POINTER(ctypes.c_double):
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------------------------------
# Options as the reference's command line names them


@dataclass
class Parsed:
    method: int
    options: OracleOptions
    data: list[OracleData]
    errors: list[str] = field(default_factory=list)


def _items(text: str) -> list[str]:
    return text.split(",")


def parse_options(arguments: list[str], components: int) -> Parsed:
    """The oracle's structures of options given as '--name value' or '--name=value'."""
    given: dict[str, str] = {}
    flags: set[str] = set()
    index = 0
    while index < len(arguments):
        argument = arguments[index]
        name, _, value = argument[2:].partition("=")
        if name in ("no-weight-scaling", "no-momentum", "warnings-are-errors"):
            flags.add(name)
        elif "=" in argument:
            given[name] = value
        else:
            index += 1
            given[name] = arguments[index]
        index += 1
    methods = {"mmse": 0, "tv": 1, "tgv": 2}
    method = methods[given.get("method", "tgv")]
    o = OracleOptions()
    o.method = method
    o.alpha = float(given.get("alpha", 1.0))
    o.alpha1 = float(given.get("alpha1", 1.0))
    o.alpha0 = float(given.get("alpha0", 2.0))
    o.coupled = 0 if given.get("channels", "coupled") == "apart" else 1
    if "channel-weights" in given:
        weights = [float(item) for item in _items(given["channel-weights"])]
        o.has_channel_weights = 1
        for c, value in enumerate(weights[:3]):
            o.channel_weights[c] = value
    if "iterations" in given:
        o.has_iterations, o.iterations = 1, int(given["iterations"])
    if "tolerance" in given:
        o.has_tolerance, o.tolerance = 1, float(given["tolerance"])
    o.relative_tolerance = float(given.get("relative-tolerance", 0.0))
    o.partial_tolerance = float(given.get("partial-tolerance", 0.0))
    if "partial-radius" in given:
        o.has_partial_radius, o.partial_radius = 1, float(given["partial-radius"])
    if "step-ratio" in given:
        o.has_step_ratio, o.step_ratio = 1, float(given["step-ratio"])
    if "relaxation" in given:
        o.has_relaxation, o.relaxation = 1, float(given["relaxation"])
    o.step_product = float(given.get("step-product", 0.99))
    if "norm-squared" in given:
        o.has_norm_squared, o.norm_squared = 1, float(given["norm-squared"])
    o.weight_scaling = 0 if "no-weight-scaling" in flags else 1
    o.record_every = int(given.get("record-every", 10))
    o.free_radius = float(given.get("free-radius", 255.0))

    def pick(name: str, default: str, c: int) -> str:
        if name not in given:
            return default
        items = _items(given[name])
        return items[0] if len(items) == 1 else items[c]

    data = []
    for c in range(components):
        d = OracleData()
        mu = pick("mu", "rule", c)
        d.has_mu, d.mu = (0, 0.0) if mu == "rule" else (1, float(mu))
        d.mu_scale = float(pick("mu-scale", "9.0", c))
        d.mu_power = float(pick("mu-power", "0.9", c))
        d.mu_chroma = float(pick("mu-chroma", "0.3", c))
        d.slack = float(pick("slack", "0.0", c))
        d.slack_cost = float(pick("slack-cost", "0.0", c))
        d.dc_weight = float(pick("dc-weight", "1.0", c))
        d.power = float(pick("weight-power", "2.0", c))
        d.flat_weight = float(pick("flat-weight", "0.1", c))
        d.flat_levels = int(pick("flat-levels", "0", c))
        flat_dc = pick("flat-dc", "1", c)
        d.has_flat_dc, d.flat_dc = (0, 0) if flat_dc == "any" else (1, int(flat_dc))
        d.centres = 0 if pick("centres", "midpoint", c) == "mmse" else 1
        data.append(d)
    return Parsed(method, o, data)


# ---------------------------------------------------------------------------------------
# Solving


@dataclass
class Component:
    levels: np.ndarray  # (rows, columns, 8, 8) int16
    table: np.ndarray  # (64,) uint16, natural order
    h: int = 1
    v: int = 1


@dataclass
class OracleResult:
    canvas: np.ndarray  # (C, H, W)
    coefficients: list[np.ndarray]
    lower: list[np.ndarray]
    upper: list[np.ndarray]
    centres: list[np.ndarray]
    steps: list[np.ndarray]
    weights: list[np.ndarray]
    flat_weights: list[np.ndarray]
    flat: list[np.ndarray]
    cells: list[tuple[int, int]]
    tau: float
    sigma: float
    iterations: int
    stop: int
    record_iterations: np.ndarray
    primal: np.ndarray
    dual: np.ndarray
    theta: np.ndarray
    partial: np.ndarray
    iteration_seconds: float = 0.0
    least_iteration_seconds: float = 0.0
    iterate: np.ndarray | None = None  # the last x, (C, H, W); canvas is the last x~


def solve(height: int, width: int, color_space: int, components: list[Component], options: list[str],
          start: dict | None = None) -> OracleResult:
    """Runs the oracle. start may hold 'coefficients' (list per component), 'p' (C,2,H,W),
    'w' (C,2,H,W) and 'r' (C,3,H,W)."""
    lib = library()
    count = len(components)
    parsed = parse_options(options, count)
    raw = (OracleComponent * count)()
    keep = []
    for index, component in enumerate(components):
        levels = np.ascontiguousarray(component.levels, dtype=np.int16)
        keep.append(levels)
        raw[index].levels = levels.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))
        raw[index].rows, raw[index].columns = int(levels.shape[0]), int(levels.shape[1])
        raw[index].h, raw[index].v = int(component.h), int(component.v)
        for k, step in enumerate(np.asarray(component.table).reshape(64)):
            raw[index].table[k] = int(step)
    data = (OracleData * count)(*parsed.data)
    start_struct = None
    if start is not None:
        start_struct = OracleStart()
        for c, coefficients in enumerate(start.get("coefficients") or []):
            array = np.ascontiguousarray(coefficients, dtype=np.float64)
            keep.append(array)
            start_struct.coefficients[c] = _doubles(array)
        for name in ("p", "w", "r"):
            if start.get(name) is not None:
                array = np.ascontiguousarray(start[name], dtype=np.float64)
                keep.append(array)
                setattr(start_struct, name, _doubles(array))
    message = ctypes.create_string_buffer(1024)
    handle = lib.oracle_solve(height, width, color_space, count, raw, data, ctypes.byref(parsed.options),
                              ctypes.byref(start_struct) if start_struct is not None else None, message, 1024)
    if not handle:
        raise ValueError(message.value.decode())
    try:
        C, H, W = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        pointer = lib.oracle_result_canvas(handle, ctypes.byref(C), ctypes.byref(H), ctypes.byref(W))
        shape = (C.value, H.value, W.value)
        canvas = np.ctypeslib.as_array(pointer, shape=(int(np.prod(shape)),)).copy().reshape(shape)
        iterate = np.ctypeslib.as_array(lib.oracle_result_iterate(handle), shape=(int(np.prod(shape)),)).copy()
        arrays: list[list[np.ndarray]] = [[] for _ in range(8)]
        cells = []
        for c in range(count):
            br, bc, down, across = (ctypes.c_uint64() for _ in range(4))
            lib.oracle_result_component(handle, c, ctypes.byref(br), ctypes.byref(bc), ctypes.byref(down),
                                        ctypes.byref(across))
            cells.append((down.value, across.value))
            for which in range(8):
                n = ctypes.c_uint64()
                p = lib.oracle_result_array(handle, c, which, ctypes.byref(n))
                values = np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0)
                if which <= 3:
                    values = values.reshape(br.value, bc.value, 8, 8)
                elif which <= 6:
                    values = values.reshape(8, 8)
                else:
                    values = values.reshape(br.value, bc.value)
                arrays[which].append(values)
        tau, sigma = ctypes.c_double(), ctypes.c_double()
        iterations, stop = ctypes.c_uint64(), ctypes.c_int32()
        lib.oracle_result_solver(handle, ctypes.byref(tau), ctypes.byref(sigma), ctypes.byref(iterations),
                                 ctypes.byref(stop))
        n = ctypes.c_uint64()
        p = lib.oracle_result_record_iterations(handle, ctypes.byref(n))
        record_iterations = np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint64)
        values = []
        for which in range(4):
            p = lib.oracle_result_record_values(handle, which, ctypes.byref(n))
            values.append(np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0))
        return OracleResult(
            canvas=canvas, coefficients=arrays[0], lower=arrays[1], upper=arrays[2], centres=arrays[3],
            steps=arrays[4], weights=arrays[5], flat_weights=arrays[6], flat=arrays[7], cells=cells,
            tau=tau.value, sigma=sigma.value, iterations=iterations.value, stop=stop.value,
            record_iterations=record_iterations, primal=values[0], dual=values[1], theta=values[2],
            partial=values[3], iteration_seconds=float(lib.oracle_result_iteration_seconds(handle)),
            least_iteration_seconds=float(lib.oracle_result_least_iteration_seconds(handle)),
            iterate=iterate.reshape(shape),
        )
    finally:
        lib.oracle_result_free(handle)


def set_threads(count: int) -> None:
    """The threads the oracle's iterations take (0: every core); results do not depend on it."""
    library().oracle_set_threads(int(count))


def threads() -> int:
    return int(library().oracle_threads())


def basis() -> np.ndarray:
    out = np.zeros(64)
    library().oracle_basis(_doubles(out))
    return out.reshape(8, 8)


def dct_forward(block: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(block, dtype=np.float64).reshape(64)
    out = np.zeros(64)
    library().oracle_dct_forward(_doubles(src), _doubles(out))
    return out.reshape(8, 8)


def dct_inverse(block: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(block, dtype=np.float64).reshape(64)
    out = np.zeros(64)
    library().oracle_dct_inverse(_doubles(src), _doubles(out))
    return out.reshape(8, 8)


def laplace_scale(zeros: int, others: int, total: int, step: float) -> float:
    return float(library().oracle_laplace_scale(zeros, others, total, step))


def laplace_shrinkage(rho: float) -> float:
    return float(library().oracle_laplace_shrinkage(rho))


def mmse_centre(level: int, step: float, scale: float) -> float:
    return float(library().oracle_mmse_centre(level, step, scale))


def sum_rows(terms: np.ndarray) -> float:
    array = np.ascontiguousarray(terms, dtype=np.float64)
    rows, columns = array.shape
    return float(library().oracle_sum_rows(_doubles(array), rows, columns))


def colour_space_code(name: str) -> int:
    return {"grayscale": 1, "ycbcr": 2, "rgb": 3}[name]


# Pictures of rounded samples (docs/math.md, 11).

def _levels16(levels) -> np.ndarray:
    return np.ascontiguousarray(levels, dtype=np.uint16)


def samples_intervals(levels, bits: int, slack: float = 0.0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The lower and upper ends and the centres of levels of `bits` bits."""
    values = _levels16(levels)
    low, high, centre = (np.zeros(values.size) for _ in range(3))
    made = library().oracle_samples_intervals(values.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), values.size,
                                              bits, slack, _doubles(low), _doubles(high), _doubles(centre))
    if not made:
        raise ValueError("bits, a level or the slack refused")
    return low.reshape(values.shape), high.reshape(values.shape), centre.reshape(values.shape)


def samples_prox(levels, bits: int, values, tau: float, mu: float = 0.0, slack: float = 0.0) -> np.ndarray:
    given = _levels16(levels)
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros(given.size)
    library().oracle_samples_prox(given.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), given.size, bits, mu, slack,
                                  _doubles(v), tau, _doubles(out))
    return out.reshape(given.shape)


def _samples_sum(name: str, levels, bits: int, mu: float, slack: float, field) -> float:
    given = _levels16(levels)
    values = np.ascontiguousarray(field, dtype=np.float64)
    channels, height, width = given.shape
    return float(getattr(library(), name)(given.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), channels, height,
                                          width, bits, mu, slack, _doubles(values)))


def samples_value(levels, bits: int, x, mu: float = 0.0, slack: float = 0.0) -> float:
    """G(x) of a (C, H, W) picture of levels."""
    return _samples_sum("oracle_samples_value", levels, bits, mu, slack, x)


def samples_conjugate(levels, bits: int, xi, mu: float = 0.0, slack: float = 0.0) -> float:
    """G*(xi) of a (C, H, W) picture of levels."""
    return _samples_sum("oracle_samples_conjugate", levels, bits, mu, slack, xi)


@dataclass
class SamplesResult:
    picture: np.ndarray  # (C, H, W)
    tau: float
    sigma: float
    iterations: int
    stop: int
    record_iterations: np.ndarray
    primal: np.ndarray
    dual: np.ndarray
    theta: np.ndarray
    partial: np.ndarray
    least_iteration_seconds: float = 0.0


def solve_samples(levels, bits: int, options: list[str]) -> SamplesResult:
    """TV or TGV within the intervals of a (C, H, W) picture of levels; --mu and --slack of
    the options are the data term's."""
    lib = library()
    given = _levels16(levels)
    channels, height, width = given.shape
    parsed = parse_options(options, channels)
    term = parsed.data[0]
    mu = term.mu if term.has_mu else 0.0
    message = ctypes.create_string_buffer(1024)
    handle = lib.oracle_samples_solve(given.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), channels, height, width,
                                      bits, mu, term.slack, ctypes.byref(parsed.options), message, 1024)
    if not handle:
        raise ValueError(message.value.decode())
    try:
        C, H, W = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        pointer = lib.oracle_result_canvas(handle, ctypes.byref(C), ctypes.byref(H), ctypes.byref(W))
        shape = (C.value, H.value, W.value)
        picture = np.ctypeslib.as_array(pointer, shape=(int(np.prod(shape)),)).copy().reshape(shape)
        tau, sigma = ctypes.c_double(), ctypes.c_double()
        iterations, stop = ctypes.c_uint64(), ctypes.c_int32()
        lib.oracle_result_solver(handle, ctypes.byref(tau), ctypes.byref(sigma), ctypes.byref(iterations),
                                 ctypes.byref(stop))
        n = ctypes.c_uint64()
        p = lib.oracle_result_record_iterations(handle, ctypes.byref(n))
        record_iterations = np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint64)
        values = []
        for which in range(4):
            p = lib.oracle_result_record_values(handle, which, ctypes.byref(n))
            values.append(np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0))
        return SamplesResult(picture, tau.value, sigma.value, iterations.value, stop.value, record_iterations,
                             *values, least_iteration_seconds=float(lib.oracle_result_least_iteration_seconds(handle)))
    finally:
        lib.oracle_result_free(handle)
