"""Writes tests/golden/blue_noise.json: a digest of the reference's blue-noise matrix.

The reference keeps the 64 x 64 ranks of its ordered dither (docs/math.md, 10) in
rust/unround/src/blue_noise.rs, which its scripts/blue_noise.py generates by
Ulichney's void-and-cluster method. This library generates the matrix itself; to pin
that generation without keeping the reference's table, this script keeps only the
SHA-256 of the ranks (row by row, each a little-endian uint16), their count and sum,
and the places of the lowest ranks. Run it from the repository's root:

    python tests/golden/make_blue_noise_digest.py [/root/reference/rust/unround/src/blue_noise.rs]
"""

import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np

SOURCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/rust/unround/src/blue_noise.rs")
TARGET = Path(__file__).resolve().parent / "blue_noise.json"


def main() -> None:
    text = SOURCE.read_text()
    body = text[text.index("= [") + 3 : text.index("];")]
    ranks = np.array([int(value) for value in re.findall(r"\d+", body)], dtype="<u2")
    if ranks.size != 4096 or sorted(ranks.tolist()) != list(range(4096)):
        raise SystemExit(f"{SOURCE} does not hold 4096 ranks, each of 0 to 4095 once")
    digest = {
        "source": "rust/unround/src/blue_noise.rs (RANKS, row by row)",
        "count": int(ranks.size),
        "sum": int(ranks.astype(np.int64).sum()),
        "sha256_u16le": hashlib.sha256(ranks.tobytes()).hexdigest(),
        "places_of_ranks_0_to_7": [int(np.flatnonzero(ranks == rank)[0]) for rank in range(8)],
    }
    TARGET.write_text(json.dumps(digest, indent=2) + "\n")
    print(f"wrote {TARGET}")


if __name__ == "__main__":
    main()
