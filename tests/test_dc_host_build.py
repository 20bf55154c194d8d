"""The tridiagonal divide and conquer's numerics (imago_amd/csrc/
tridiag_dc.hpp: deflation, secular roots, Loewner vectors, rotations,
merge tree) compiled for the host and driven merge by merge: eigenvalues
against LAPACK, orthogonality and residual at rounding level on every
case of tests/tridiag_cases.py (no GPU needed)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import tridiag_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_dc_numerics_host_build(tmp_path):
    exe = tmp_path / "check_dc"
    src = os.path.join(ROOT, "tests", "host_build", "check_dc.cpp")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-host-only", src,
                    "-o", str(exe)], check=True, capture_output=True, timeout=600)
    cs = tridiag_cases.cases(max_n=300)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(struct.pack("i", len(cs)))
        for _, d, e in cs:
            f.write(struct.pack("i", len(d)))
            f.write(np.asarray(d, float).tobytes())
            f.write(np.asarray(e, float).tobytes())
    subprocess.run([str(exe), str(fin), str(fout)], check=True, timeout=300)
    raw = np.fromfile(fout)
    o = 0
    for name, d, e in cs:
        n = len(d)
        w = raw[o:o + n]
        o += n
        Z = raw[o:o + n * n].reshape(n, n).T
        o += n * n
        ev, orth, res = tridiag_cases.check(d, e, w, Z)
        print("%-16s n=%4d eig %.1e orth %.1e res %.1e" % (name, n, ev, orth, res))
        assert np.all(np.diff(w) >= 0), name
        assert ev < 1e-13 and orth < 1e-13 and res < 1e-13, name
    assert o == raw.size
