"""The conformance cases of the reference as fixtures (tests/golden/conformance/*.npz).

Each case is a small quantized picture, the options of its reconstruction, and what the
reference made of it, with tolerances derived from rounding bounds
(conformance/README.md:27-45 states the pass criteria that tests/test_oracle_conformance.py
and tests/test_gpu_conformance.py apply).
"""

from dataclasses import dataclass
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "conformance"
SPACES = {"grayscale": 1, "ycbcr": 2, "rgb": 3}
U = 2.0**-53


@dataclass
class Case:
    name: str
    height: int
    width: int
    space: int
    components: list  # (h, v, rows, columns)
    tables: list
    levels: list
    options: list[str]
    data: dict

    @property
    def method(self) -> str:
        words = self.options
        return words[words.index("--method") + 1]

    def start(self) -> dict | None:
        if "first-coefficients_0" not in self.data:
            return None
        C = len(self.components)
        coefficients = [self.data[f"first-coefficients_{c}"] for c in range(C)]
        H, W = self.canvas().shape[1:]

        def field(name: str, entries: int):
            if f"{name}_0_0" not in self.data:
                return None
            return np.stack([np.stack([self.data[f"{name}_{c}_{e}"] for e in range(entries)]) for c in range(C)])

        start = {"coefficients": coefficients, "p": field("first-p", 2), "w": field("first-w", 2),
                 "r": field("first-r", 3)}
        for name in ("p", "w", "r"):
            if start[name] is not None:
                assert start[name].shape[2:] == (H, W)
        return start

    def canvas(self) -> np.ndarray:
        C = len(self.components)
        return np.stack([self.data[f"canvas_{c}"] for c in range(C)])


def names() -> list[str]:
    return sorted(path.stem for path in GOLDEN.glob("*.npz"))


def load(name: str) -> Case:
    data = dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
    height, width = (int(v) for v in data["picture"])
    components = [tuple(int(v) for v in row) for row in data["components"]]
    C = len(components)
    return Case(
        name=name,
        height=height,
        width=width,
        space=SPACES[str(data["space"])],
        components=components,
        tables=[data[f"table_{c}"] for c in range(C)],
        levels=[data[f"levels_{c}"] for c in range(C)],
        options=str(data["options"]).split(),
        data=data,
    )


def centres_within(centres: np.ndarray, case: Case, c: int) -> bool:
    """Every centre within (2|q| + 20) u Q of its exact value where it is an MMSE one, and
    that value itself elsewhere (conformance/README.md:33-36)."""
    hi, lo = case.data[f"exact-hi_{c}"], case.data[f"exact-lo_{c}"]
    levels = case.levels[c][: hi.shape[0], : hi.shape[1]].astype(np.float64)
    steps = case.tables[c].astype(np.float64).reshape(8, 8)
    distance = np.abs((centres - hi) - lo)
    bound = (2.0 * np.abs(levels) + 20.0) * U * steps
    words = case.options
    centres_option = words[words.index("--centres") + 1] if "--centres" in words else "midpoint"
    mmse = centres_option == "mmse" or case.method == "mmse"
    ac = np.ones((8, 8), bool)
    ac[0, 0] = False
    inexact = mmse & ac[None, None] & (levels != 0)
    return bool(np.all(np.where(inexact, distance <= bound, centres == hi)))
