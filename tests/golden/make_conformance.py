"""Writes the conformance fixtures under tests/golden/conformance/ from the reference's cases.

The reference's cases (conformance/cases/*.txt, format in conformance/README.md:47-74)
are data: the quantized coefficients of small pictures, the options of their
reconstruction, and what the reference made of them with its tolerances. This script
reads them from the reference's checkout and keeps the numbers, one .npz per case, so
that the tests need no access to the reference. Run it from the repository's root:

    python tests/golden/make_conformance.py [/root/reference/conformance/cases]
"""

import sys
from pathlib import Path

import numpy as np

SOURCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/conformance/cases")
TARGET = Path(__file__).resolve().parent / "conformance"


def numbers(values: list[str]) -> np.ndarray:
    return np.array([float(v) for v in values], dtype=np.float64)


def convert(path: Path) -> dict:
    out: dict = {}
    components: dict[int, tuple[int, int, int, int]] = {}
    rows: dict[str, dict] = {}
    records, tolerance_records = [], []
    for line in path.read_text().splitlines():
        if not line or line.startswith("#"):
            continue
        words = line.split()
        key = words[0]
        if key == "name":
            out["name"] = words[1]
        elif key == "picture":
            out["picture"] = np.array([int(words[1]), int(words[2])])
            out["space"] = words[3]
        elif key == "component":
            components[int(words[1])] = tuple(int(w) for w in words[2:6])
        elif key == "table":
            out[f"table_{words[1]}"] = np.array([int(w) for w in words[2:]], dtype=np.uint16)
        elif key == "options":
            out["options"] = " ".join(words[1:])
        elif key == "steps":
            out["steps"] = numbers(words[1:])
        elif key == "weights":
            out[f"weights_{words[1]}"] = numbers(words[2:])
        elif key in ("levels", "lower", "upper", "centres", "first-coefficients", "canvas"):
            rows.setdefault(f"{key}_{words[1]}", {})[int(words[2])] = words[3:]
        elif key == "exact-centres":
            pairs = [w.split(":") for w in words[3:]]
            rows.setdefault(f"exact-hi_{words[1]}", {})[int(words[2])] = [p[0] for p in pairs]
            rows.setdefault(f"exact-lo_{words[1]}", {})[int(words[2])] = [p[1] for p in pairs]
        elif key in ("first-p", "first-w", "first-r"):
            # entry, channel, row
            rows.setdefault(f"{key}_{words[2]}_{words[1]}", {})[int(words[3])] = words[4:]
        elif key == "record":
            records.append([float(w) for w in words[1:6]])
        elif key == "stop":
            out["stop"] = np.array([int(words[1]), 1 if words[2] == "converged" else 2])
        elif key == "tolerance":
            if words[1] == "canvas":
                out["tolerance_canvas"] = float(words[2])
            elif words[1] == "record":
                tolerance_records.append([float(w) for w in words[2:6]])
    out["components"] = np.array([components[k] for k in sorted(components)], dtype=np.int64)
    for name, by_row in rows.items():
        kind = name.split("_")[0]
        stacked = [by_row[k] for k in sorted(by_row)]
        if kind == "levels":
            array = np.array([[int(w) for w in row] for row in stacked], dtype=np.int16)
            out[name] = array.reshape(len(stacked), -1, 8, 8)
        elif kind in ("lower", "upper", "centres", "first-coefficients", "exact-hi", "exact-lo"):
            array = np.array([[float(w) for w in row] for row in stacked], dtype=np.float64)
            out[name] = array.reshape(len(stacked), -1, 8, 8)
        else:
            out[name] = np.array([[float(w) for w in row] for row in stacked], dtype=np.float64)
    out["records"] = np.array(records, dtype=np.float64).reshape(-1, 5)
    out["tolerance_records"] = np.array(tolerance_records, dtype=np.float64).reshape(-1, 4)
    return out


def main() -> None:
    TARGET.mkdir(parents=True, exist_ok=True)
    for path in sorted(SOURCE.glob("*.txt")):
        case = convert(path)
        np.savez_compressed(TARGET / f"{path.stem}.npz", **case)
        print(path.stem, sorted(case)[:6], "...")


if __name__ == "__main__":
    main()
