"""The oracle's data term at the defaults that every benchmarked configuration runs with,
pinned to the reference's own vectors:

* the default constants and the rule of mu (cpp/tests/test_model.cpp:161-177): no mu
  given, scale 9, power 0.9, chroma factor 0.3, DC weight 1, midpoint centres, weight
  power 2, flat weight 0.1, flat levels 0, flat DC bound 1; mu = 9 pow(mean step, 0.9)
  for luma and (9 x 0.3) pow(mean step, 0.9) for the chroma of YCbCr, the platform's pow
  (the same libm serves the oracle and this test);
* flat blocks (cpp/tests/test_model.cpp:441-530): the marks by the count of AC levels
  that are not 0 for flat_levels 0, 1, 2, 62 and 63, and by the DC neighbours' bound for
  the DC levels 0 1 5 / 0 3 5 and bounds none, 0, 1, 2, 3, 4; a flat block's weights are
  the flat weight times the others', exactly;
* rust/unround/tests/model.rs:296-312: with flat_levels 63 every block is flat.

The conformance cases pass --flat-weight 1.0 and an explicit mu, so these are what anchor
the default path. Each check of a vector is also run on a planted mistake (the oracle
given another option) and must fail on it."""

import math

import numpy as np

import helpers
from oracle import oracle
from unround import jpegio, synthetic


def problem(levels: np.ndarray, table: np.ndarray, options: list[str]) -> oracle.OracleResult:
    """The oracle's model of a grey picture's levels (rows, columns, 8, 8), no iteration."""
    rows, columns = levels.shape[:2]
    component = oracle.Component(levels.astype(np.int16), table.reshape(64).astype(np.uint16))
    return oracle.solve(8 * rows, 8 * columns, int(jpegio.ColorSpace.GRAYSCALE), [component],
                        ["--method", "tv", "--iterations", "0"] + options)


def ac_count_levels() -> np.ndarray:
    # Block 0 has DC alone, block 1 one AC level, block 2 two, block 3 all 63
    # (cpp/tests/test_model.cpp:442-453); entries in natural order, row by row.
    levels = np.zeros((2, 2, 64), dtype=np.int16)
    levels[0, 0, 0] = 5
    levels[0, 1, 0] = -3
    levels[0, 1, 9] = 1
    levels[1, 0, 1] = 2
    levels[1, 0, 63] = -1
    levels[1, 1, :] = 1
    return levels.reshape(2, 2, 8, 8)


AC_CASES = {0: [1, 0, 0, 0], 1: [1, 1, 0, 0], 2: [1, 1, 1, 0], 62: [1, 1, 1, 0], 63: [1, 1, 1, 1]}


def test_flat_blocks_by_their_ac_levels():
    levels = ac_count_levels()
    table = np.full(64, 10, dtype=np.uint16)
    for most, expected in AC_CASES.items():
        made = problem(levels, table, ["--mu", "0.5", "--flat-weight", "0.25", "--flat-levels", str(most),
                                       "--flat-dc", "any"])
        assert made.flat[0].reshape(4).astype(int).tolist() == expected, most
        # A flat block's weights are a quarter of the others', exactly, the DC's among them.
        assert np.array_equal(made.flat_weights[0], 0.25 * made.weights[0])
    # Planted: a bound one off marks other blocks.
    planted = problem(levels, table, ["--mu", "0.5", "--flat-levels", "1", "--flat-dc", "any"])
    assert planted.flat[0].reshape(4).astype(int).tolist() != AC_CASES[0]


def dc_bound_levels() -> np.ndarray:
    # Blocks of 2 x 3 with no AC level, DC levels 0 1 5 / 0 3 5 (test_model.cpp:495-501).
    levels = np.zeros((2, 3, 8, 8), dtype=np.int16)
    for index, dc in enumerate((0, 1, 5, 0, 3, 5)):
        levels[index // 3, index % 3, 0, 0] = dc
    return levels


DC_CASES = {"any": [1, 1, 1, 1, 1, 1], "0": [0, 0, 0, 0, 0, 0], "1": [1, 0, 0, 0, 0, 0], "2": [1, 0, 0, 0, 0, 1],
            "3": [1, 0, 0, 1, 1, 1], "4": [1, 1, 1, 1, 1, 1]}


def test_flat_blocks_by_their_dc_neighbours():
    levels = dc_bound_levels()
    table = np.full(64, 10, dtype=np.uint16)
    for bound, expected in DC_CASES.items():
        made = problem(levels, table, ["--flat-weight", "0.5", "--flat-dc", bound])
        assert made.flat[0].reshape(6).astype(int).tolist() == expected, bound
    # Planted: the bound of 2 read as 3 marks other blocks.
    assert problem(levels, table, ["--flat-weight", "0.5", "--flat-dc", "3"]).flat[0].reshape(6).astype(
        int).tolist() != DC_CASES["2"]


def test_every_block_is_flat_at_63_levels():
    rng = np.random.default_rng(5)
    levels = rng.integers(-20, 21, size=(3, 4, 8, 8)).astype(np.int16)
    levels[..., 0, 0] = rng.integers(-100, 100, size=(3, 4))
    table = rng.integers(1, 60, size=64).astype(np.uint16)
    made = problem(levels, table, ["--mu", "2.0", "--flat-weight", "0.25", "--flat-levels", "63", "--flat-dc", "any"])
    assert np.all(made.flat[0] == 1)
    assert np.array_equal(made.flat_weights[0], 0.25 * made.weights[0])


def test_default_constants_and_the_rule_of_mu_for_luma():
    # The 256 x 256 grey picture of seed 0 at quality 50, as the TV benchmark takes it.
    picture = synthetic.grey_picture(256, 256, 0)
    image = synthetic.encode_grey(picture, 50)
    levels = np.asarray(image.components[0].coefficients)
    table = np.asarray(image.components[0].quant_table, dtype=np.float64).reshape(64)
    made = problem(levels, table, [])
    mean = float(np.sum(table.astype(np.uint32))) / 64.0
    assert mean == 57.625  # Annex K luma at quality 50
    mu = 9.0 * math.pow(mean, 0.9)
    expected = np.array([mu / (step * step) for step in table]).reshape(8, 8)  # DC weight 1, power 2
    assert np.array_equal(made.weights[0], expected)
    assert np.array_equal(made.steps[0], table.reshape(8, 8))
    assert np.array_equal(made.flat_weights[0], made.weights[0] * 0.1)
    # Midpoint centres: q Q, the DC shifted by 1024; intervals (q -+ 1/2) Q.
    q = levels.astype(np.int64)
    steps = table.reshape(8, 8).astype(np.int64)
    shift = np.zeros((8, 8))
    shift[0, 0] = 1024.0
    assert np.array_equal(made.centres[0], (q * steps).astype(np.float64) + shift)
    assert np.array_equal(made.lower[0], ((2 * q - 1) * steps).astype(np.float64) / 2.0 + shift)
    assert np.array_equal(made.upper[0], ((2 * q + 1) * steps).astype(np.float64) / 2.0 + shift)
    # The default flat marks are those of flat_levels 0 and a DC bound of 1.
    explicit = problem(levels, table, ["--flat-levels", "0", "--flat-dc", "1", "--flat-weight", "0.1"])
    assert np.array_equal(made.flat[0], explicit.flat[0])
    assert 0 < int(np.sum(made.flat[0])) < made.flat[0].size  # some blocks are flat, some not
    # Planted: another power or scale moves the weights.
    assert not np.array_equal(problem(levels, table, ["--mu-power", "1.0"]).weights[0], expected)
    assert not np.array_equal(problem(levels, table, ["--mu-scale", "9.000000000000002"]).weights[0], expected)


def test_the_rule_of_mu_for_chroma():
    picture = synthetic.colour_picture(32, 32, 4)
    image = synthetic.encode_colour420(picture, 75)
    luma, chroma = image.components[0], image.components[1]
    made = helpers.oracle_solve(image, ["--method", "tv", "--iterations", "0"])
    for index, component in enumerate(image.components):
        table = np.asarray(component.quant_table, dtype=np.float64).reshape(64)
        mean = float(np.sum(table.astype(np.uint32))) / 64.0
        scale = 9.0 if index == 0 else 9.0 * 0.3  # the scale times the chroma factor first
        mu = scale * math.pow(mean, 0.9)
        expected = np.array([mu / (step * step) for step in table]).reshape(8, 8)
        assert np.array_equal(made.weights[index], expected), index
    assert np.asarray(luma.quant_table).sum() != np.asarray(chroma.quant_table).sum()
    # Planted: 9 x pow x 0.3 (the factor last) is another double for these tables.
    table = np.asarray(chroma.quant_table, dtype=np.float64).reshape(64)
    mean = float(np.sum(table.astype(np.uint32))) / 64.0
    reordered = 9.0 * math.pow(mean, 0.9) * 0.3
    order_matters = reordered != (9.0 * 0.3) * math.pow(mean, 0.9)
    if order_matters:
        assert not np.array_equal(made.weights[1], np.array([reordered / (s * s) for s in table]).reshape(8, 8))
    # An explicit mu is taken as it is, in the chroma too.
    given = helpers.oracle_solve(image, ["--method", "tv", "--iterations", "0", "--mu", "2.5"])
    assert given.weights[1][0, 1] == 2.5 / (table[1] * table[1])
