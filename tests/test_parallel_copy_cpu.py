"""The fork-join copy group that fills the page-locked staging ring
(vship_amd/csrc/parallel_copy.hpp), built with g++ and run on the CPU:
part ranges tile every copy exactly and copies equal memcpy for 1-9
threads, with no byte written past the end."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(not shutil.which("g++"), reason="g++ not available")
def test_parallel_copy_matches_memcpy(tmp_path):
    exe = tmp_path / "parallel_copy_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-Wall", "-Werror", "-o", str(exe),
                    os.path.join(ROOT, "tests", "parallel_copy_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
