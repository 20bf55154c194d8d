"""The eigensolver's divide-and-conquer failure codes mapped onto the
k-points of the half that failed (map_stedc_info in
imago_amd/csrc/hetrd_launch.hpp), compiled for the host: a code in 1..n
(zheevd's "did not converge") for exactly the matrices whose rows the
LAPACK dstedc info names, 0 for the others."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_stedc_info_mapping(tmp_path):
    exe = tmp_path / "check_stedc_info"
    src = os.path.join(ROOT, "tests", "host_build", "check_stedc_info.cpp")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-host-only", src,
                    "-o", str(exe)], check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(" ok") == 9
