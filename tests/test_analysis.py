"""Symbolic analysis checks on the host (no GPU).

The threaded block symbolic factorization (one elimination-tree height at a
time) must give the fronts, parents and levels of the serial block symbolic
factorization over the same ordering. `analysis_check.cpp` recomputes them
serially from the lower pattern and compares. The batch instances 12, 30
and 62 (seed 1000) have poses whose odometry partners sit in disjoint
dissection subtrees, so their supernodal elimination tree is not the
dissection tree.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sdf-slam_amd", "lib")


@pytest.fixture(scope="module")
def checker(sdfslam, tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("analysis") / "analysis_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "sdf-slam_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "sdf-slam_amd", "host"),
                    os.path.join(ROOT, "tests", "analysis_check.cpp"),
                    "-L" + LIB, "-lsdfslam_host", "-lsdfslam_hip", "-Wl,-rpath," + LIB,
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("seed,names", [
    (1000, ["batch12", "batch30", "batch62", "batch3", "batch29"]),
    (1, ["corner", "room_small", "office_small", "room"]),
])
def test_threaded_symbolic_matches_serial(checker, seed, names):
    out = subprocess.run([checker, str(seed)] + names, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.strip().splitlines()
    assert len(lines) == len(names) and all(line.endswith(" ok") for line in lines), out.stdout
