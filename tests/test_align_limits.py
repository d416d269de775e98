"""The traceback lane kernel's packed end-cell key (32 bits, (row - i_first)
* W + k): the host-side limit that routes jobs of too many rows to the
one-thread kernel, compiled and checked on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_lane_key_limit(tmp_path):
    exe = str(tmp_path / "align_limits_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(HERE, "align_limits_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout
