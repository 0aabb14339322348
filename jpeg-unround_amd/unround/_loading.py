"""Where the library is: at the path in UNROUND_LIBRARY, or in unround/_native
(python/src/unround/_loading.py:77-101 of the reference, for its library's name)."""

import os
from pathlib import Path
from typing import Final

__all__ = ["LIBRARY", "bundled", "configured", "file_name"]

LIBRARY: Final = "unround_capi"
"""The stem of the C interface's library (include/unround.h), built by jpeg-unround_amd/build.py."""


def file_name(stem: str) -> str:
    return f"lib{stem}.so"


def configured(variable: str) -> Path | None:
    value = os.environ.get(variable)
    return Path(value) if value else None


def bundled(stem: str) -> Path | None:
    path = Path(__file__).parent / "_native" / file_name(stem)
    return path if path.is_file() else None
