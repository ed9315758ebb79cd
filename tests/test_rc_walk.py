"""CPU replay of the row x column kernels' walks against the planner's
parameters (tests/c/rc_walk_check.cpp, compiled here with g++ together with
the planner source): for the lane-fixed walk, LDS and table bounds, exact
coverage of every tile element with its true A and B offsets, and
out-of-range offsets for every padding slot; for the general walk on a
sample of launches, full and partial tiles, that the staging and read
positions agree element by element and stay in bounds; and lane-fixed launches
whose A offsets pass 2^30 (no padding slots on A's side)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_lane_fixed_walk_index_arithmetic(tmp_path):
    exe = str(tmp_path / "rc_walk_check")
    src = [os.path.join(ROOT, "tests", "c", "rc_walk_check.cpp"), os.path.join(ROOT, "ttc_amd", "csrc", "ttc_plan.cpp")]
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "ttc_amd", "csrc")] + src + ["-o", exe]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        pytest.fail("g++ failed:\n" + p.stdout)
    r = subprocess.run([exe, "300"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    line = r.stdout.strip().splitlines()[-1]
    assert line.endswith(": ok"), line
    checked, fast, general, big = [int(w) for w in line.replace(",", " ").split() if w.isdigit()][:4]
    assert checked > 10000 and fast > checked // 40 and general > checked // 10 and big > 0, line
