"""The CPU oracle against the reference's conformance cases (conformance/README.md:27-45).

Intervals, weights and steps agree to the last bit; centres within (2|q| + 20) u Q of
their exact values; the canvas within the case's tolerance; every record's primal value
within its tolerance, the dual values certified against the other's primal value, and
the gaps within their tolerance where the case gives one; the same stop. The
subgradient method is not part of the path, and its case is left out.
"""

import numpy as np
import pytest

import cases
from oracle import oracle

PDHG_CASES = [name for name in cases.names() if name != "grey-subgradient"]


def run(case: cases.Case) -> oracle.OracleResult:
    components = [
        oracle.Component(levels=case.levels[c], table=case.tables[c], h=case.components[c][0],
                         v=case.components[c][1])
        for c in range(len(case.components))
    ]
    return oracle.solve(case.height, case.width, case.space, components, case.options, case.start())


@pytest.mark.parametrize("name", PDHG_CASES)
def test_oracle_reproduces_case(name):
    case = cases.load(name)
    result = run(case)
    C = len(case.components)
    for c in range(C):
        assert np.array_equal(result.lower[c], case.data[f"lower_{c}"]), "lower ends"
        assert np.array_equal(result.upper[c], case.data[f"upper_{c}"]), "upper ends"
        assert np.array_equal(result.weights[c].reshape(64), case.data[f"weights_{c}"]), "weights"
        assert cases.centres_within(result.centres[c], case, c), "centres"
    if "steps" in case.data:
        assert result.tau == case.data["steps"][0]
        assert result.sigma == case.data["steps"][1]
    expected = case.canvas()
    distance = float(np.sqrt(np.sum((result.canvas - expected) ** 2)))
    assert distance <= case.data["tolerance_canvas"], f"canvas {distance} > {case.data['tolerance_canvas']}"
    records = case.data["records"]
    tolerances = case.data["tolerance_records"]
    assert list(result.record_iterations) == [int(r) for r in records[:, 0]]
    for k, (row, tol) in enumerate(zip(records, tolerances, strict=True)):
        assert int(tol[0]) == int(row[0])
        primal, dual = result.primal[k], result.dual[k]
        assert abs(primal - row[1]) <= tol[1], f"record {k} primal {primal} vs {row[1]}"
        if tol[2] != -1.0:
            assert dual <= row[1] + tol[2]
            assert row[2] <= primal + tol[2]
        if tol[3] != -1.0:
            assert abs((primal - dual) - (row[1] - row[2])) <= tol[3]
    stop_iterations, reason = (int(v) for v in case.data["stop"])
    assert result.iterations == stop_iterations
    if case.method in ("tv", "tgv"):
        assert result.stop == reason
