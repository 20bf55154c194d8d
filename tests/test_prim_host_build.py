"""The device primitive evaluators (imago_amd/csrc/prim.hpp) compiled for
the host and checked against the oracle on every (l_a, l_b) class: a CPU
check of the GPU code's mathematics (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_device_primitives_host_build_match_oracle(tmp_path, oracle_mod):
    exe = tmp_path / "check_prims"
    src = os.path.join(ROOT, "tests", "host_build", "check_prims.cpp")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-host-only", src,
                    "-o", str(exe), "-ldl"], check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe), os.path.join(ROOT, "oracle", "liboracle.so")],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
