"""scfV.dat reader/writer (initPotCoeffs, potential.f90:433-564; writer
potentialUpdate.F90:1672-1722)."""
import numpy as np
import pytest

from imago_amd import configs
from imago_amd.scfv import ScfvMismatch, _fortran_e, read_scfv, write_scfv


def test_fortran_e_format():
    assert _fortran_e(12.5, 17, 10) == " 0.1250000000E+02"
    assert _fortran_e(-0.00123, 13, 6) == "-0.123000E-02"
    assert _fortran_e(0.0, 13, 6) == " 0.000000E+00"
    assert len(_fortran_e(-9.99999999999, 17, 10)) == 17


def test_round_trip(tmp_path):
    s = configs.si2_primitive()
    c = configs.pot_coefficients(s, seed=1)
    p = tmp_path / "scfV.dat"
    write_scfv(p, s, c)
    got = read_scfv(p, s)[:, 0]
    assert np.allclose(got, c, rtol=1e-10)
    text = p.read_text().splitlines()
    assert text[0] == "NUM_TYPES    1" and text[1] == "TOTAL__OR__SPIN_UP"


def test_term_count_mismatch_is_an_error(tmp_path):
    s = configs.si2_primitive()
    s_other = configs.si2_primitive(npot=12)
    p = tmp_path / "scfV.dat"
    write_scfv(p, s_other, np.zeros(s_other.pot_dim))
    with pytest.raises(ScfvMismatch):
        read_scfv(p, s)
