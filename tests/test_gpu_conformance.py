"""The GPU's binary64 path (--precision double) against the reference's conformance cases,
through the C interface, by the reference's own criteria (conformance/README.md:27-45):
intervals, weights and steps to the last bit; centres within (2|q| + 20) u Q; the last
tilde canvas (the last record's, as an observer sees it) within the case's tolerance; the
records within theirs; the same stop. The result's canvas, projected so that it
requantizes, lies within the tolerance plus what that projection may move (2^-30 a
coefficient). The subgradient case is not part of the path."""

import numpy as np
import pytest

import cases
import unround
from unround import native

pytestmark = pytest.mark.gpu

PDHG_CASES = [name for name in cases.names() if name != "grey-subgradient"]


def image_of(case: cases.Case) -> unround.jpegio.Image:
    components = tuple(
        unround.jpegio.make_component(case.levels[c], case.tables[c], case.components[c][0], case.components[c][1])
        for c in range(len(case.components))
    )
    return unround.jpegio.make_image(case.height, case.width, unround.jpegio.ColorSpace(case.space), components)


@pytest.mark.parametrize("name", PDHG_CASES)
def test_gpu_double_reproduces_case(name):
    case = cases.load(name)
    print(f"conformance case {name}", flush=True)
    seen = []
    decoded = native.solve(image_of(case), case.options + ["--precision", "double"], observer=seen.append,
                           start=case.start())
    C = len(case.components)
    for c in range(C):
        problem = decoded.problems[c]
        assert np.array_equal(problem.lower, case.data[f"lower_{c}"])
        assert np.array_equal(problem.upper, case.data[f"upper_{c}"])
        assert np.array_equal(problem.weights.reshape(64), case.data[f"weights_{c}"])
        assert cases.centres_within(problem.centres, case, c)
        coefficients = decoded.coefficients[c]
        assert np.all(coefficients > problem.lower) and np.all(coefficients < problem.upper)
    expected = case.canvas()
    tolerance = case.data["tolerance_canvas"]
    # The observer sees every record after the first, the stopping one too (as the
    # reference's run loop behind its C interface calls it, rust/unround/src/pdhg.rs:645-665).
    if seen and decoded.result is not None and seen[-1].iteration == decoded.result.iterations:
        raw = seen[-1].canvas
        distance = float(np.sqrt(np.sum((raw - expected) ** 2)))
        assert distance <= tolerance, f"last tilde canvas {distance} > {tolerance}"
    moved = 2.0**-30 * np.sqrt(sum(64 * case.levels[c][:, :].shape[0] * case.levels[c].shape[1] for c in range(C)))
    final = float(np.sqrt(np.sum((decoded.canvas - expected) ** 2)))
    assert final <= tolerance + moved
    history = decoded.result.history if decoded.result is not None else None
    records, tolerances = case.data["records"], case.data["tolerance_records"]
    if history is not None:
        assert list(history.iterations) == [int(r) for r in records[:, 0]]
        for k, (row, tol) in enumerate(zip(records, tolerances, strict=True)):
            primal, dual = history.primal[k], history.dual[k]
            assert abs(primal - row[1]) <= tol[1], f"record {k}: primal {primal} vs {row[1]}"
            if tol[2] != -1.0:
                assert dual <= row[1] + tol[2] and row[2] <= primal + tol[2]
            if tol[3] != -1.0:
                assert abs((primal - dual) - (row[1] - row[2])) <= tol[3]
            if not np.isnan(row[3]):
                assert abs(history.scaling[k] - row[3]) <= 1e-9
            if not np.isnan(row[4]):
                assert abs(history.partial_gap[k] - row[4]) <= tol[1] + 1e-6 * abs(row[4])
        stop_iterations, reason = (int(v) for v in case.data["stop"])
        assert decoded.result.iterations == stop_iterations
        assert int(decoded.result.stop) == reason
