"""The interleaved-row index maps of the structured-grid engine (ir_pos, ir_x,
ir_shift in lbmminiapp_amd/csrc/lbm_dense_kernels.hpp), compiled for the host
and checked on the CPU against their stated definitions for rows of 64 K
cells, K = 1 .. 16:

* ir_pos is a bijection of [0, nx) and ir_x is its inverse;
* ir_shift(pos(x), d) = pos((x + d) mod nx) for d = -1, 0, 1 (periodic);
* the property the odd step relies on: the 64 positions of block w map, for
  d = +-1, onto the 64 positions of block (w + d) mod K, so a wavefront that
  covers one block reads and writes one whole aligned block.

A planted mistake (a non-periodic shift at the row ends) fails the check."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHECK = r'''
#include <cstdio>
#include <vector>
#include "lbm_dense_kernels.hpp"
using namespace lbm;
#ifdef PLANT_MISTAKE
static int shift(int p, int d, int K) {  // plain +-1 on positions: not periodic at the ends
  int q = p + (d > 0 ? 64 : (d < 0 ? -64 : 0));
  return q < 0 ? 0 : (q >= 64 * K ? 64 * K - 1 : q);
}
#else
static int shift(int p, int d, int K) { return ir_shift(p, d, K); }
#endif
int main() {
  long bad = 0;
  for (int K = 1; K <= 16; K++) {
    const int nx = 64 * K;
    std::vector<int> seen(nx, 0);
    for (int x = 0; x < nx; x++) {
      const int p = ir_pos(x, K);
      if (p < 0 || p >= nx || ir_x(p, K) != x) bad++;
      else seen[p]++;
    }
    for (int p = 0; p < nx; p++) bad += seen[p] != 1;
    for (int p = 0; p < nx; p++)
      for (int d = -1; d <= 1; d++) {
        const int q = shift(p, d, K);
        if (q < 0 || q >= nx || ir_x(q, K) != ((ir_x(p, K) + d) % nx + nx) % nx) bad++;
      }
    for (int w = 0; w < K; w++)
      for (int d = -1; d <= 1; d += 2) {
        std::vector<int> hit(64, 0);
        for (int l = 0; l < 64; l++) {
          const int q = shift(64 * w + l, d, K);
          if (q / 64 != ((w + d) % K + K) % K) bad++;
          else hit[q % 64]++;
        }
        for (int l = 0; l < 64; l++) bad += hit[l] != 1;
      }
  }
  std::printf("%ld\n", bad);
  return bad != 0;
}
'''


def _build_and_run(tmp_path, plant):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc")
                                      else None)
    if hipcc is None:
        pytest.skip("no hipcc")
    src = tmp_path / "irmap.cpp"
    src.write_text(CHECK)
    exe = tmp_path / "irmap"
    cmd = [hipcc, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "lbmminiapp_amd", "csrc"), str(src), "-o", str(exe)]
    if plant:
        cmd.insert(1, "-DPLANT_MISTAKE")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)


def test_interleaved_row_maps(tmp_path):
    r = _build_and_run(tmp_path, plant=False)
    assert r.returncode == 0 and r.stdout.strip() == "0", (r.stdout, r.stderr)


def test_interleaved_row_maps_catch_a_planted_mistake(tmp_path):
    r = _build_and_run(tmp_path, plant=True)
    assert r.returncode != 0 and int(r.stdout.strip()) > 0
