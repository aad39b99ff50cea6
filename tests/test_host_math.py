"""CPU: rvs_math.h (shared by the host library and the kernels) emulates the
reference's double-accumulated cv::norm bit for bit, and valid_tri's maximum of
three such norms, without FP64; checked against real double arithmetic on
random, midpoint-steered and special inputs."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


import pytest


@pytest.mark.parametrize("faithful", [False, True])
def test_norm_emulation_matches_double(tmp_path, faithful):
    """faithful: square root and reciprocal perturbed by up to one ulp, as the
    device's hardware instructions guarantee."""
    exe = str(tmp_path / "test_rvs_math")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off"] + (["-DRVS_FAITHFUL"] if faithful else []) +
                   ["-o", exe, os.path.join(ROOT, "tests", "cpp", "test_rvs_math.cpp")], check=True)
    r = subprocess.run([exe, "300000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fast mismatches 0, slow mismatches 0" in r.stdout
    assert " 0 mismatches" in r.stdout   # triangle stretch (valid_tri's max of three double norms)
