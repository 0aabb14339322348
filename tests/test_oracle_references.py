"""The oracle's parts against the reference's exact references (tests/golden/references.npz).

The rules are those of the reference's own tests (rust/unround/tests/references.rs:32-156):
every basis entry is the double nearest to it; the DCT lies within its rounding bound
(docs/math.md:108-122) of the 60-digit transform; the Laplace scale and the shrinkage
within 16 units of 2^-53; a centre within (2|q| + 17) u Q of the mean of its bin. Each
check is shown to fail on a planted mistake (a transposed transform, a dropped term).
"""

from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle

REFERENCES = dict(np.load(Path(__file__).resolve().parent / "golden" / "references.npz", allow_pickle=False))
U = 2.0**-53


def gamma(k: int) -> float:
    return k * U / (1 - k * U)


def within(value: float, hi: float, lo: float, bound: float) -> bool:
    return abs(Fraction(value) - Fraction(hi) - Fraction(lo)) <= Fraction(bound)


def test_basis_is_the_nearest_doubles():
    basis = oracle.basis()
    for (k, n), hi in zip(REFERENCES["basis_index"], REFERENCES["basis_hi"], strict=True):
        assert basis[k, n] == hi


def dct_within(computed: np.ndarray, kind: int, block: np.ndarray, hi: np.ndarray, lo: np.ndarray) -> bool:
    C = np.abs(oracle.basis())
    X = np.abs(block.astype(np.float64).reshape(8, 8))
    magnitude = C @ X @ C.T if kind == 0 else C.T @ X @ C
    epsilon = gamma(6) * (2 + gamma(6))
    bound = 1.0001 * epsilon * magnitude
    flat = computed.reshape(64)
    return all(within(flat[i], hi[i], lo[i], bound.reshape(64)[i]) for i in range(64))


@pytest.mark.parametrize("index", range(8))
def test_dct_within_its_rounding(index):
    kind = int(REFERENCES["dct_kind"][index])
    block = REFERENCES["dct_block"][index]
    transform = oracle.dct_forward if kind == 0 else oracle.dct_inverse
    computed = transform(block.reshape(8, 8).astype(np.float64))
    assert dct_within(computed, kind, block, REFERENCES["dct_hi"][index], REFERENCES["dct_lo"][index])
    # A transposed output (the row and column passes' roles swapped) fails the check, and
    # where the block is symmetric, a relative error of 1e-12 does.
    symmetric = np.array_equal(block.reshape(8, 8), block.reshape(8, 8).T)
    planted = computed * (1 + 1e-12) if symmetric else computed.T
    assert not dct_within(planted, kind, block, REFERENCES["dct_hi"][index], REFERENCES["dct_lo"][index])


def test_dct_round_trip_and_orthogonality():
    rng = np.random.default_rng(7)
    block = rng.integers(-128, 128, size=(8, 8)).astype(np.float64)
    assert np.max(np.abs(oracle.dct_inverse(oracle.dct_forward(block)) - block)) < 1e-12
    basis = oracle.basis()
    assert np.max(np.abs(basis @ basis.T - np.eye(8))) < 4 * U


def test_scale_within_sixteen_units():
    for counts, hi, lo in zip(REFERENCES["scale_counts"], REFERENCES["scale_hi"], REFERENCES["scale_lo"],
                              strict=True):
        zeros, others, total, step = (int(v) for v in counts)
        scale = oracle.laplace_scale(zeros, others, total, float(step))
        assert within(scale, hi, lo, 16 * U * abs(hi)), (counts, scale)


def test_shrinkage_within_sixteen_units():
    previous = 0.0
    for index, (rho, hi, lo) in enumerate(
        zip(REFERENCES["shrinkage_rho"], REFERENCES["shrinkage_hi"], REFERENCES["shrinkage_lo"], strict=True)
    ):
        value = oracle.laplace_shrinkage(float(rho))
        assert within(value, hi, lo, 16 * U), (rho, value)
        if index < 3001:
            assert value >= previous
            previous = value


def test_centres_are_the_means_of_their_bins():
    for (level, step, scale), hi, lo in zip(REFERENCES["centres_given"], REFERENCES["centres_hi"],
                                            REFERENCES["centres_lo"], strict=True):
        centre = oracle.mmse_centre(int(level), float(step), float(scale))
        bound = (2 * abs(level) + 17) * U * step
        assert within(centre, hi, lo, bound), (level, step, scale, centre)
        assert (abs(level) - 0.5) * step <= abs(centre) <= abs(level) * step
        assert (centre < 0) == (level < 0)


def test_sums_of_rows_are_exact_on_dyadics_and_fixed_in_order():
    rng = np.random.default_rng(3)
    terms = rng.integers(-2**20, 2**20, size=(300, 37)).astype(np.float64) / 1024.0
    assert oracle.sum_rows(terms) == float(sum(Fraction(v) for v in terms.reshape(-1)))
    # Eight lanes, then rows in blocks of 128 and a tree: a term placed 1e16 away from its
    # cancelling partner in another lane is lost or kept by that order alone.
    row = np.zeros((1, 16))
    row[0, 0], row[0, 8], row[0, 1] = 1e16, -1e16, 1.0
    assert oracle.sum_rows(row) == 1.0
