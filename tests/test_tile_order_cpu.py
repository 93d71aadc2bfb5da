"""The workgroup -> tile order of the int8 products (fgls_tile_of_block,
fgls_tile_grid in fgls_device.h), compiled for the host and run on the CPU:
for every XCD group shape the kernels use, each (row tile, column tile) is
produced by exactly one workgroup, padding workgroups produce none, and the
tiles of one group all land on one XCD (workgroup b runs on XCD b % 8)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include <vector>
#include "fgls_device.h"

template <int GR, int GC>
int check(int RT, int CT) {
  const int G = fgls_tile_grid(RT, CT, GR, GC);
  std::vector<int> seen((size_t)RT * CT, 0), xcd((size_t)RT * CT, -1), grp((size_t)RT * CT, -1);
  for (int b = 0; b < G; ++b) {
    int rt = -1, ct = -1;
    if (!fgls_tile_of_block<GR, GC>(b, RT, CT, rt, ct)) continue;
    if (rt < 0 || rt >= RT || ct < 0 || ct >= CT) { std::printf("out of range %d %d\n", rt, ct); return 1; }
    const size_t t = (size_t)rt * CT + ct;
    seen[t] += 1;
    xcd[t] = b & 7;
    grp[t] = (b >> 3) / (GR * GC);
  }
  for (size_t t = 0; t < seen.size(); ++t)
    if (seen[t] != 1) { std::printf("tile %zu seen %d times (GR %d GC %d RT %d CT %d)\n", t, seen[t], GR, GC, RT, CT); return 1; }
  // the tiles of one group (same row group and column group) share an XCD
  for (int rt = 0; rt < RT; ++rt)
    for (int ct = 0; ct < CT; ++ct) {
      const size_t t = (size_t)rt * CT + ct, t0 = (size_t)(rt / GR * GR) * CT + ct / GC * GC;
      if (xcd[t] != xcd[t0] || grp[t] != grp[t0]) { std::printf("group split (GR %d GC %d)\n", GR, GC); return 1; }
    }
  return 0;
}

int main() {
  const int RTs[] = {1, 7, 79, 313, 938}, CTs[] = {1, 3, 64, 256, 257};
  for (int RT : RTs)
    for (int CT : CTs) {
      if (check<16, 4>(RT, CT) || check<16, 8>(RT, CT) || check<32, 2>(RT, CT) || check<8, 8>(RT, CT)) return 1;
    }
  std::printf("OK\n");
  return 0;
}
"""


def test_tile_order_is_a_bijection_per_group_shape(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    src = tmp_path / "tile_order.hip"
    src.write_text(PROG)
    exe = tmp_path / "tile_order"
    subprocess.run([hipcc, "-O1", "-std=c++17", "--offload-arch=gfx950", "-I",
                    os.path.join(ROOT, "cugwas-proto_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.strip() == "OK", res.stdout + res.stderr
