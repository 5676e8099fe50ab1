"""The kernels' block-index division (FastDiv in csrc/lbm_common.hpp) is exact (CPU).

Compiles a host-only check with hipcc against the same header and compares
(n * m) >> k with n / d for edge values and random n < 2^31 over many divisors.
"""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

SRC = r'''
#include <cstdio>
#include <random>
#include "lbm_common.hpp"
int main() {
  std::mt19937_64 g(7);
  long bad = 0, n_tests = 0;
  const uint32_t ds[] = {1, 2, 3, 5, 7, 16, 17, 255, 256, 257, 4095, 4096, 4097, 65535, 65536,
                         123457, (1u << 30) - 1, (1u << 30) + 1, 2147483647u};
  for (uint32_t d : ds) {
    plbm::FastDiv f = plbm::make_fastdiv(d);
    const uint32_t edge[] = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 2147483647u, 2147483646u};
    for (uint32_t n : edge) {
      if (n >= 2147483648u) continue;
      ++n_tests; bad += f.div(n) != n / d;
    }
    for (int t = 0; t < 100000; ++t) {
      uint32_t n = g() & 0x7fffffffu; ++n_tests; bad += f.div(n) != n / d;
    }
  }
  for (int t = 0; t < 1000; ++t) {
    uint32_t d = (uint32_t)(g() & 0x7fffffffu) | 1u;
    plbm::FastDiv f = plbm::make_fastdiv(d);
    for (int u = 0; u < 500; ++u) {
      uint32_t n = g() & 0x7fffffffu; ++n_tests; bad += f.div(n) != n / d;
    }
  }
  printf("%ld %ld\n", n_tests, bad);
  return bad != 0;
}
'''


def test_fastdiv_exact(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    src = tmp_path / "fastdiv.cpp"
    src.write_text(SRC)
    exe = tmp_path / "fastdiv"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-I", str(ROOT / "pylbm_amd" / "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    n_tests, bad = map(int, out.stdout.split())
    assert out.returncode == 0 and bad == 0 and n_tests > 2_000_000
