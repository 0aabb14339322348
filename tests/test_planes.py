"""The choice between buffer-resource and 64-bit addressing (csrc/planes.hpp), on the CPU
with no allocation: a lane with nothing to do addresses 2^31 bytes into its resource, which
the hardware's range check turns into a zero load or a dropped store only where the
resource spans at most 2^31 bytes. So every kind of field a kernel takes through buffer
resources must span at most that, in the prox step, the dual step, the batch's and the
unrolled solve's alike. The header is compiled here by itself (it has no HIP in it)."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER_DIR = ROOT / "jpeg-unround_amd" / "csrc"

PROGRAM = r"""
#include <cstdio>
#include "planes.hpp"
using namespace unround::planes;
int main() {
  std::printf("limit %zu outside %u\n", kBufferSpanLimit, kOutsideOffset);
  const size_t cases[][5] = {
      // channels, samples, t, s, dual (prox is printed beside it)
      {1, 178956970, 4, 4, 0}, {1, 178956971, 4, 4, 0}, {1, 300000000, 4, 4, 0},
      {1, 268435456, 4, 2, 0}, {1, 268435457, 4, 2, 0},
      {3, 59652323, 4, 4, 0}, {3, 59652324, 4, 4, 0},
      {1, 89478485, 8, 8, 0}, {1, 89478486, 8, 8, 0},
      {1, 16777216, 4, 4, 0}, {3, 4194304, 4, 4, 0}};
  for (const auto& c : cases) {
    std::printf("%zu %zu %zu %zu dual %d prox %d\n", c[0], c[1], c[2], c[3], dual_fits(c[0], c[1], c[2], c[3]) ? 1 : 0,
                prox_fits(c[1], c[2], c[3]) ? 1 : 0);
  }
  std::printf("fits %d %d\n", fits(kBufferSpanLimit) ? 1 : 0, fits(kBufferSpanLimit + 1) ? 1 : 0);
  std::printf("forced %d buffers %d %d\n", wide_forced() ? 1 : 0, dual_buffers(1, 1024, 4, 4) ? 1 : 0,
              prox_buffers(1024, 4, 4) ? 1 : 0);
  return 0;
}
"""


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    compiler = shutil.which("g++") or shutil.which("c++")
    if compiler is None:
        pytest.skip("no host C++ compiler")
    folder = tmp_path_factory.mktemp("planes")
    source = folder / "planes_check.cpp"
    source.write_text(PROGRAM)
    binary = folder / "planes_check"
    subprocess.run([compiler, "-std=c++17", "-O1", f"-I{HEADER_DIR}", str(source), "-o", str(binary)], check=True)
    return binary


def run(binary, forced: bool) -> dict:
    env = dict(os.environ)
    env.pop("UNROUND_WIDE_PLANES", None)
    if forced:
        env["UNROUND_WIDE_PLANES"] = "1"
    lines = subprocess.run([str(binary)], env=env, check=True, capture_output=True, text=True).stdout.splitlines()
    table = {}
    for line in lines[1:-2]:
        words = line.split()
        table[tuple(int(w) for w in words[:4])] = (int(words[5]), int(words[7]))
    return {"head": lines[0], "table": table, "fits": lines[-2], "forced": lines[-1]}


def test_the_sentinel_is_the_span_limit(program):
    out = run(program, False)
    assert out["head"] == f"limit {2**31} outside {2**31}"
    assert out["fits"] == "fits 1 0"


def test_dual_and_prox_bounds_at_the_edges(program):
    table = run(program, False)["table"]
    # binary32 grey TGV: r spans 3 x 4 N bytes, so N <= floor(2^31 / 12) = 178956970.
    assert table[(1, 178956970, 4, 4)][0] == 1
    assert table[(1, 178956971, 4, 4)][0] == 0
    # Between 2 and 4 GiB: the span a bound of 4 GiB would have let through.
    assert table[(1, 300000000, 4, 4)] == (0, 0)
    # BF16 duals: w (2 x 4 N) is the widest kind, N <= 2^28.
    assert table[(1, 268435456, 4, 2)][0] == 1
    assert table[(1, 268435457, 4, 2)][0] == 0
    # Three channels: r spans 9 x 4 N.
    assert table[(3, 59652323, 4, 4)][0] == 1
    assert table[(3, 59652324, 4, 4)][0] == 0
    # binary64: r spans 3 x 8 N; the prox's p 2 x 8 N.
    assert table[(1, 89478485, 8, 8)] == (1, 1)
    assert table[(1, 89478486, 8, 8)][0] == 0
    # The benchmarked canvases stay on buffer resources: 4096 x 4096 grey, 3 x 2048 x 2048.
    assert table[(1, 16777216, 4, 4)] == (1, 1)
    assert table[(3, 4194304, 4, 4)] == (1, 1)


def test_the_environment_forces_wide_planes(program):
    assert run(program, False)["forced"] == "forced 0 buffers 1 1"
    assert run(program, True)["forced"] == "forced 1 buffers 0 0"
