"""Writes tests/golden/references.npz from the reference's exact references.

conformance/references/*.txt (written by the reference from 60-70 digit decimals) hold
the DCT basis, blocks and their transforms, Laplace scales, shrinkages and centres, each
exact value as a pair of doubles hi:lo. This keeps those numbers as arrays. Run from the
repository's root:

    python tests/golden/make_references.py [/root/reference/conformance/references]
"""

import sys
from pathlib import Path

import numpy as np

SOURCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/conformance/references")
TARGET = Path(__file__).resolve().parent / "references.npz"


def rows(name: str) -> list[list[str]]:
    text = (SOURCE / name).read_text()
    return [line.split() for line in text.splitlines() if line.strip() and not line.startswith("#")]


def pairs(words: list[str]) -> tuple[np.ndarray, np.ndarray]:
    split = [w.split(":") for w in words]
    return np.array([float(a) for a, _ in split]), np.array([float(b) for _, b in split])


def main() -> None:
    out = {}
    basis = rows("basis.txt")
    hi, lo = pairs([r[2] for r in basis])
    out["basis_index"] = np.array([[int(r[0]), int(r[1])] for r in basis])
    out["basis_hi"], out["basis_lo"] = hi, lo
    dct = rows("dct.txt")
    kinds, blocks, his, los = [], [], [], []
    for given, exact in zip(dct[0::2], dct[1::2], strict=True):
        assert exact[0] == "exact"
        kinds.append(0 if given[0] == "forward" else 1)
        blocks.append([int(w) for w in given[1:]])
        h, l = pairs(exact[1:])
        his.append(h)
        los.append(l)
    out["dct_kind"], out["dct_block"] = np.array(kinds), np.array(blocks)
    out["dct_hi"], out["dct_lo"] = np.array(his), np.array(los)
    scale = rows("scale.txt")
    out["scale_counts"] = np.array([[int(w) for w in r[:4]] for r in scale], dtype=np.uint64)
    out["scale_hi"], out["scale_lo"] = pairs([r[4] for r in scale])
    shrinkage = rows("shrinkage.txt")
    out["shrinkage_rho"] = np.array([float(r[0]) for r in shrinkage])
    out["shrinkage_hi"], out["shrinkage_lo"] = pairs([r[1] for r in shrinkage])
    centres = rows("centres.txt")
    out["centres_given"] = np.array([[float(w) for w in r[:3]] for r in centres])
    out["centres_hi"], out["centres_lo"] = pairs([r[3] for r in centres])
    np.savez_compressed(TARGET, **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
