"""The oracle's iterations in bands of rows over threads (OpenMP, static bands): every
sample is made by one thread in the same operations whatever the count, so a solve gives
the same bits on 1, 3 and 8 threads, as the reference's solves do across its threads'
strips (cpp/tests/test_pdhg.cpp:328-386); the records (fixed-order sums) too."""

import numpy as np
import pytest

import helpers
from oracle import oracle
from unround import synthetic


@pytest.fixture
def restore_threads():
    before = oracle.threads()
    yield
    oracle.set_threads(before)


@pytest.mark.parametrize("make,options", [
    (lambda: synthetic.encode_colour420(synthetic.colour_picture(72, 88, 7), 75),
     ["--method", "tgv", "--iterations", "12", "--tolerance", "0", "--record-every", "5"]),
    (lambda: synthetic.encode_grey(synthetic.grey_picture(64, 72, 8), 50),
     ["--method", "tv", "--iterations", "12", "--tolerance", "0", "--record-every", "4"]),
])
def test_thread_counts_give_the_same_bits(restore_threads, make, options):
    image = make()
    results = []
    for count in (1, 3, 8):
        oracle.set_threads(count)
        results.append(helpers.oracle_solve(image, options))
    first = results[0]
    for other in results[1:]:
        assert np.array_equal(first.canvas, other.canvas)
        assert np.array_equal(first.primal, other.primal) and np.array_equal(first.dual, other.dual)
        for a, b in zip(first.coefficients, other.coefficients):
            assert np.array_equal(a, b)
    assert first.iteration_seconds > 0.0
