"""Builds the library: hipcc for gfx950, every source of csrc/ into
unround/_native/libunround_capi.so (the name the reference's Python package looks for).

    python jpeg-unround_amd/build.py
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SOURCES = HERE / "csrc"
OBJECTS = HERE / "_build"
TARGET = HERE / "unround" / "_native" / "libunround_capi.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Host-only sources: the ROCm clang as a plain C++ compiler.
CXX = os.environ.get("UNROUND_CXX", "/opt/rocm/lib/llvm/bin/clang++")
ARCH = os.environ.get("UNROUND_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-pthread", "-std=c++20", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", f"-I{SOURCES}"]
# Device code: gfx950 only; packed FP32 instructions off (they are made by the compiler
# from plain float code, and are kept out on purpose).
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def compile_one(source: Path) -> Path:
    out = OBJECTS / (source.stem + ".o")
    if source.suffix == ".hip":
        command = [HIPCC, "-x", "hip", *DEVICE, *COMMON, "-c", str(source), "-o", str(out)]
    else:
        command = [CXX, "-x", "c++", *COMMON, "-c", str(source), "-o", str(out)]
    subprocess.run(command, check=True)
    return out


def build() -> Path:
    OBJECTS.mkdir(exist_ok=True)
    TARGET.parent.mkdir(parents=True, exist_ok=True)
    sources = sorted(SOURCES.glob("*.hip")) + sorted(SOURCES.glob("*.cpp"))
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objects = list(pool.map(compile_one, sources))
    temporary = TARGET.with_suffix(".so.tmp")
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(temporary), *map(str, objects), "-lz", "-pthread"],
                   check=True)
    temporary.replace(TARGET)
    return TARGET


if __name__ == "__main__":
    print(build())
    sys.exit(0)
