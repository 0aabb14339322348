"""The library's data term at the defaults equals the oracle's bit for bit: the weights
(mu by the rule of the mean step, power 0.9, the chroma's factor 0.3; DC weight 1;
divided by the squared steps), the steps, the interval ends and the midpoint centres of
every component, through unround_result_problem. The oracle's defaults are pinned to the
reference's vectors by tests/test_oracle_defaults.py, so the library's are pinned through
it; the flat marks enter each solve, which test_gpu_parity.py compares with the oracle at
the defaults (the 256 x 256 TV solve at quality 50 has 111 flat blocks of 1024)."""

import numpy as np
import pytest

import helpers
from unround import native, synthetic

pytestmark = pytest.mark.gpu

CASES = [
    ("grey 256x256 q50", lambda: synthetic.encode_grey(synthetic.grey_picture(256, 256, 0), 50)),
    ("grey 512x512 q30", lambda: synthetic.encode_grey(synthetic.grey_picture(512, 512, 1), 30)),
    ("4:2:0 128x96 q75", lambda: synthetic.encode_colour420(synthetic.colour_picture(128, 96, 0), 75)),
]


@pytest.mark.parametrize("label,make", CASES, ids=[case[0] for case in CASES])
def test_default_problem_equals_the_oracles(label, make):
    image = make()
    options = ["--method", "tgv", "--iterations", "1", "--tolerance", "0"]
    print(f"defaults: {label}", flush=True)
    ours = native.solve(image, options)
    theirs = helpers.oracle_solve(image, ["--method", "tgv", "--iterations", "0"])
    for c, problem in enumerate(ours.problems):
        rows, columns = problem.lower.shape[:2]
        assert np.array_equal(problem.weights, theirs.weights[c]), f"component {c}: weights"
        assert np.array_equal(problem.steps, theirs.steps[c]), f"component {c}: steps"
        assert np.array_equal(problem.lower, theirs.lower[c][:rows, :columns]), f"component {c}: lower"
        assert np.array_equal(problem.upper, theirs.upper[c][:rows, :columns]), f"component {c}: upper"
        assert np.array_equal(problem.centres, theirs.centres[c][:rows, :columns]), f"component {c}: centres"
