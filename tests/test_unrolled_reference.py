"""The unrolled TGV restatement (unrolled_reference.py, binary64 torch) against the oracle:
with the weight map m = 1 and the method's alpha0 = 2, alpha1 = 1, the unrolled iteration
is the method's own PDHG iteration (docs/math.md:544-566 with one radius), so N iterations
from the oracle's start and problem must reach the oracle's last iterate x (not its
canvas, the last x~), to binary64 rounding. This pins the restatement the GPU's unrolled
forward and adjoint are checked against."""

import numpy as np
import pytest
import torch

import helpers
import unrolled_reference
from unround import synthetic


@pytest.mark.parametrize("size,seed,quality,iterations", [(64, 1, 30, 12), (48, 2, 75, 7)])
def test_unrolled_restatement_is_the_method(size, seed, quality, iterations):
    image = synthetic.encode_grey(synthetic.grey_picture(size, size, seed), quality)
    result = helpers.oracle_solve(image, ["--method", "tgv", "--iterations", str(iterations), "--tolerance", "0",
                                          "--record-every", "0"])
    start, problem = helpers.oracle_unrolled_inputs(result)
    constants = {"tau": result.tau, "sigma": result.sigma, "gamma": 1.0, "rho": 1.9}
    ones = torch.ones((size, size), dtype=torch.float64)
    x = unrolled_reference.unrolled(start[None], [problem], constants, torch.tensor(2.0, dtype=torch.float64),
                                    torch.tensor(1.0, dtype=torch.float64), ones, iterations)
    worst = float(np.max(np.abs(x[0].numpy() - result.iterate[0])))
    assert worst < 1e-9, worst
    # and the canvas is not the iterate: the comparison tells them apart
    assert float(np.max(np.abs(result.canvas[0] - result.iterate[0]))) > 1e-3
