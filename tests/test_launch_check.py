"""The host-side launch validator (anukulan_amd/csrc/launch_check.hpp), built
with g++ on the CPU: every kernel launch that records timing events goes
through it before it is enqueued (kernels.hip spmv_with and the panel
combine, batch_kernels.hip launch_rows), so a malformed launch or an event
pair split across dispatches is refused on the host and never reaches a
queue."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "anukulan_amd", "csrc")

DRIVER = r"""
#include "launch_check.hpp"
#include <cstdio>
using anuk::LaunchShape;
using anuk::launch_shape_error;
static LaunchShape shape(unsigned long long gx, unsigned long long bx,
                         unsigned long long gy = 1, unsigned long long by = 1) {
  LaunchShape s; s.gx = gx; s.bx = bx; s.gy = gy; s.by = by; return s;
}
int main() {
  int a = 1, b = 2;
  const void* e1 = &a; const void* e2 = &b;
  struct Case { LaunchShape s; const void* start; const void* stop; int ok; } cases[] = {
    {shape(1, 64), nullptr, nullptr, 1},
    {shape(1614, 256), e1, e2, 1},
    {shape(65535, 1024, 65535), nullptr, nullptr, 1},
    {shape(0, 256), nullptr, nullptr, 0},            // empty grid
    {shape(4, 0), nullptr, nullptr, 0},              // empty workgroup
    {shape(4, 1025), nullptr, nullptr, 0},           // over 1024 work-items
    {shape(4, 512, 1, 4), nullptr, nullptr, 0},      // 2048 work-items over two dims
    {shape(1ull << 31, 64), nullptr, nullptr, 0},    // grid.x too large
    {shape(4, 64, 1ull << 16), nullptr, nullptr, 0}, // grid.y too large
    {shape(4, 64), e1, nullptr, 0},                  // start only
    {shape(4, 64), nullptr, e2, 0},                  // stop only (split pair)
    {shape(4, 64), e1, e1, 0},                       // same event twice
  };
  int i = 0;
  for (const Case& c : cases) {
    const char* why = launch_shape_error(c.s, c.start, c.stop);
    std::printf("%d %d %s\n", i++, why == nullptr ? 1 : 0, why ? why : "ok");
    if ((why == nullptr) != (c.ok != 0)) return 1;
  }
  return 0;
}
"""


def test_launch_validator_rules(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "launch_check_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "launch_check_driver"
    subprocess.run([gxx, "-std=c++20", "-O1", "-Wall", "-Werror", f"-I{CSRC}", str(src), "-o",
                    str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 12
    assert "both timing events or neither" in lines[10]
