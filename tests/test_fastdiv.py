"""Host check of the invariant-divisor division used by the MGPCG index math
(csrc/dfm_fastdiv.hpp): compiled with g++ and compared with the hardware
divide on every divisor 1..4096, powers of two up to 2^30 and random large
divisors, at edge and random numerators below 2^31."""
import os
import shutil
import subprocess
import tempfile

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "differentiable-flowmap_amd", "csrc")

MAIN_SRC = r"""
#include "dfm_fastdiv.hpp"
#include <cstdio>
#include <random>
using namespace dfm;
static int check(uint32_t d, uint32_t n) {
  FastDiv f = make_fastdiv(d);
  if (fdiv(n, f) != n / d) { std::printf("FAIL d=%u n=%u got=%u\n", d, n, fdiv(n, f)); return 1; }
  return 0;
}
int main() {
  std::mt19937 rng(7);
  const uint32_t top = (1u << 31) - 1;
  int bad = 0;
  auto sweep = [&](uint32_t d) {
    uint32_t edges[] = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 2 * d, top, top - 1, top - d};
    for (uint32_t n : edges) if (n <= top) bad += check(d, n);
    for (int t = 0; t < 2000; ++t) bad += check(d, rng() & top);
    for (uint32_t q = 1; q < 64 && (uint64_t)q * d <= top; ++q) {  // exact multiples and their neighbours
      bad += check(d, q * d);
      bad += check(d, q * d - 1);
    }
  };
  for (uint32_t d = 1; d <= 4096; ++d) sweep(d);
  for (int e = 0; e <= 30; ++e) sweep(1u << e);
  for (int t = 0; t < 2000; ++t) sweep(1 + (rng() & top));
  std::printf(bad ? "FAILED %d\n" : "OK\n", bad);
  return bad ? 1 : 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_fastdiv_matches_hardware_division():
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "fd.cpp")
        exe = os.path.join(td, "fd")
        with open(src, "w") as f:
            f.write(MAIN_SRC)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout[-2000:]
        assert out.stdout.strip().endswith("OK")
