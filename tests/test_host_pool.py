"""The host thread pool behind the solvers' per-lane host work
(anukulan_amd/csrc/host_pool.cpp), checked on the CPU by a small C++ driver
(tests/cpp/host_pool_check.cpp) built here with g++."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "anukulan_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_pool(tmp_path):
    exe = str(tmp_path / "host_pool_check")
    subprocess.run(["g++", "-std=c++20", "-O2", "-pthread", "-I", CSRC,
                    os.path.join(HERE, "cpp", "host_pool_check.cpp"),
                    os.path.join(CSRC, "host_pool.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
