"""The Fortran host binding (imago_amd/fortran/imago_hip.f90) end to end:
a Fortran program calls solveZHEGV_hip with the argument list of the
reference's solveZHEGV (interfaces.F90:73-120) and gets the eigenvalues
of a known 2x2 Hermitian problem from the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "imago_amd", "fortran", "_bin", "hip_driver")


def test_fortran_driver_solves_through_the_library(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (imago_amd.build builds it)")
    print("[gpu] fortran driver", flush=True)
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    # the binding reports errors on unit 20, as the reference does
    unit20 = tmp_path / "fort.20"
    note = unit20.read_text() if unit20.exists() else ""
    assert r.returncode == 0, r.stdout + r.stderr + note
    assert len(r.stdout.split()) >= 3, r.stdout + r.stderr + note
    vals = [float(x) for x in r.stdout.split()[1:3]]
    assert vals == pytest.approx([1.0, 3.0], abs=1e-12)
