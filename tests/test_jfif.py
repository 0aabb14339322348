"""JFIF's conversions of the library (unround_to_rgb, unround_to_ycbcr) against the
rationals they stand for (docs/math.md 8; rust/unround/tests/colour.rs:174-210): the
result's RGB takes 701/500, 443/250, 25251/73375 and 209599/293500, each as its nearest
double, and the way back 299/1000, 587/1000, 114/1000, 250/443 and 500/701.

The constants are read out of the library's own outputs (at a unit chroma step each
output is one constant times one, exactly), and the nearest doubles come from exact
rational arithmetic (fractions), so that a constant off by one unit in the last place, or
T.871's six-digit 0.344136 and 0.714136, fails. Every other output is checked against the
exact model of the stated order of operations, with each product either rounded or fused
into its sum, as docs/math.md allows. Each check is also run on a planted mistake (a
neighbouring double) and must fail on it."""

from fractions import Fraction

import numpy as np
import pytest

from unround import native

pytestmark = []


def nearest(rational: Fraction) -> float:
    # int / int in Python is the correctly rounded quotient.
    return rational.numerator / rational.denominator


def rounded(value: Fraction) -> float:
    return nearest(value)


RGB_FROM = {
    "red_from_cr": Fraction(701, 500),
    "blue_from_cb": Fraction(443, 250),
    "green_from_cb": Fraction(25251, 73375),
    "green_from_cr": Fraction(209599, 293500),
}
YCBCR_FROM = {
    "y_from_red": Fraction(299, 1000),
    "y_from_green": Fraction(587, 1000),
    "y_from_blue": Fraction(114, 1000),
    "cb_from_blue": Fraction(250, 443),
    "cr_from_red": Fraction(500, 701),
}


def test_the_rationals_are_jfifs():
    # 2(1 - K_R), 2(1 - K_B), 2 K_B (1 - K_B) / K_G and 2 K_R (1 - K_R) / K_G with
    # K_R = 0.299, K_B = 0.114, K_G = 0.587; the forward ones their inverses.
    kr, kb, kg = YCBCR_FROM["y_from_red"], YCBCR_FROM["y_from_blue"], YCBCR_FROM["y_from_green"]
    assert kr + kg + kb == 1
    assert RGB_FROM["red_from_cr"] == 2 * (1 - kr)
    assert RGB_FROM["blue_from_cb"] == 2 * (1 - kb)
    assert RGB_FROM["green_from_cb"] == 2 * kb * (1 - kb) / kg
    assert RGB_FROM["green_from_cr"] == 2 * kr * (1 - kr) / kg
    assert YCBCR_FROM["cb_from_blue"] == 1 / RGB_FROM["blue_from_cb"]
    assert YCBCR_FROM["cr_from_red"] == 1 / RGB_FROM["red_from_cr"]


def library_constants() -> dict[str, float]:
    # Y = 0 with Cb = 129 or Cr = 129: each output is -, +, or 0 plus one constant times 1.
    planes = np.array([[[0.0, 0.0]], [[129.0, 128.0]], [[128.0, 129.0]]])
    rgb = native.to_rgb(planes)
    return {
        "blue_from_cb": rgb[0, 0, 2],
        "green_from_cb": -rgb[0, 0, 1],
        "red_from_cr": rgb[0, 1, 0],
        "green_from_cr": -rgb[0, 1, 1],
    }


def test_result_constants_are_the_nearest_doubles():
    found = library_constants()
    for name, rational in RGB_FROM.items():
        assert found[name] == nearest(rational), f"{name}: {found[name]!r} vs {nearest(rational)!r}"
    # The six-digit figures of T.871 are not what the library takes.
    assert found["green_from_cb"] != 0.344136 and found["green_from_cr"] != 0.714136


def rgb_model(y: float, cb: float, cr: float, constants: dict[str, float]) -> list[set[float]]:
    """Every R, G, B the stated order allows: R = Y + c (Cr - 128), B = Y + c (Cb - 128),
    G = (Y - c (Cb - 128)) - c (Cr - 128), each product rounded or fused."""
    b = Fraction(cb) - 128  # exact: the samples lie in [64, 256)
    r = Fraction(cr) - 128
    c = {name: Fraction(value) for name, value in constants.items()}
    red = {rounded(Fraction(y) + Fraction(rounded(c["red_from_cr"] * r))), rounded(Fraction(y) + c["red_from_cr"] * r)}
    blue = {rounded(Fraction(y) + Fraction(rounded(c["blue_from_cb"] * b))),
            rounded(Fraction(y) + c["blue_from_cb"] * b)}
    inner = {rounded(Fraction(y) - Fraction(rounded(c["green_from_cb"] * b))),
             rounded(Fraction(y) - c["green_from_cb"] * b)}
    green = set()
    for first in inner:
        green.add(rounded(Fraction(first) - Fraction(rounded(c["green_from_cr"] * r))))
        green.add(rounded(Fraction(first) - c["green_from_cr"] * r))
    return [red, green, blue]


def planes_sample(count: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0, 255, count), rng.uniform(64, 255, count), rng.uniform(64, 255, count)])


def misses(planes: np.ndarray, rgb: np.ndarray, constants: dict[str, float]) -> int:
    wrong = 0
    for k in range(planes.shape[1]):
        allowed = rgb_model(planes[0, k], planes[1, k], planes[2, k], constants)
        wrong += sum(rgb[k, e] not in allowed[e] for e in range(3))
    return wrong


def test_rgb_follows_the_stated_order_with_the_nearest_doubles():
    planes = planes_sample(400, 7)
    rgb = native.to_rgb(planes.reshape(3, 1, -1))[0]
    exact = {name: nearest(rational) for name, rational in RGB_FROM.items()}
    assert misses(planes, rgb, exact) == 0
    # A planted mistake: green_from_cr one unit in the last place up; some outputs must
    # then fall outside what the model allows.
    planted = dict(exact, green_from_cr=float(np.nextafter(exact["green_from_cr"], 1.0)))
    assert misses(planes, rgb, planted) > 0


def ycbcr_model(red: float, green: float, blue: float, constants: dict[str, float]) -> tuple[set[float], ...]:
    """Every Y, Cb, Cr the stated order allows: Y = (kR R + kG G) + kB B, Cb = 128 + k (B - Y),
    Cr = 128 + k (R - Y), each product rounded or fused into the sum that takes it (clang
    fuses the left product of a sum of two, or the right one)."""
    c = {name: Fraction(value) for name, value in constants.items()}
    pr, pg, pb = c["y_from_red"] * Fraction(red), c["y_from_green"] * Fraction(green), c["y_from_blue"] * Fraction(blue)
    inner = {rounded(Fraction(rounded(pr)) + Fraction(rounded(pg))), rounded(pr + Fraction(rounded(pg))),
             rounded(Fraction(rounded(pr)) + pg)}
    ys = set()
    for first in inner:
        ys.add(rounded(Fraction(first) + Fraction(rounded(pb))))
        ys.add(rounded(Fraction(first) + pb))
    cbs, crs = set(), set()
    for y in ys:
        db = Fraction(rounded(Fraction(blue) - Fraction(y)))
        dr = Fraction(rounded(Fraction(red) - Fraction(y)))
        cbs |= {rounded(128 + Fraction(rounded(c["cb_from_blue"] * db))), rounded(128 + c["cb_from_blue"] * db)}
        crs |= {rounded(128 + Fraction(rounded(c["cr_from_red"] * dr))), rounded(128 + c["cr_from_red"] * dr)}
    return ys, cbs, crs


def ycbcr_misses(picture: np.ndarray, planes: np.ndarray, constants: dict[str, float]) -> int:
    wrong = 0
    for k in range(picture.shape[0]):
        allowed = ycbcr_model(picture[k, 0], picture[k, 1], picture[k, 2], constants)
        wrong += sum(planes[e, k] not in allowed[e] for e in range(3))
    return wrong


def test_ycbcr_follows_the_stated_order_with_the_nearest_doubles():
    rng = np.random.default_rng(8)
    picture = rng.uniform(0, 255, (400, 3))
    planes = native.to_ycbcr(picture.reshape(1, -1, 3))[:, 0, :]
    exact = {name: nearest(rational) for name, rational in YCBCR_FROM.items()}
    assert ycbcr_misses(picture, planes, exact) == 0
    planted = dict(exact, cb_from_blue=float(np.nextafter(exact["cb_from_blue"], 1.0)))
    assert ycbcr_misses(picture, planes, planted) > 0


def test_sizes_that_do_not_fit_are_refused():
    library = native._lib()
    out = np.zeros(3)
    pointer = out.ctypes.data_as(native._C_TYPES["f64*"])
    assert library.unround_to_rgb(pointer, 2**63, 4, pointer) == native.Status.ARGUMENT
    assert library.unround_to_ycbcr(pointer, 2**40, 2**40, pointer) == native.Status.ARGUMENT
    assert library.unround_to_rgb(None, 1, 1, pointer) == native.Status.ARGUMENT
    with pytest.raises(ValueError):
        native.to_rgb(np.zeros((2, 4, 4)))
