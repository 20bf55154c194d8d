"""scfV.dat (fort.8): the potential-coefficient file, the boundary through
which an initial guess enters the SCF.

* ``read_scfv``   follows initPotCoeffs (src/imago/potential.f90:433-564):
                  a NUM_TYPES header, then per spin a tag line and per
                  potential type a term count followed by one line
                  ``coeff alpha rhoT rhoV rhoS`` per term, then the +U/J
                  block.  A term-count mismatch is an error, as the
                  reference's "Mismatch between scfV and imago.dat" stop.
* ``write_scfv``  follows the writer in makeSCFPot
                  (src/imago/potentialUpdate.F90:1672-1722): formats
                  (a9,i5), (i5), (2(1x,e17.10),3(1x,e13.6)), and the
                  NUM__PLUSUJ__TERMS block (no +U/J atoms here).

Coefficient order is potential-type order, the order of the potential
terms j used by the library (types are numbered in the order of their
first site).
"""
from __future__ import annotations

import numpy as np


class ScfvMismatch(ValueError):
    """scfV.dat disagrees with the system's potential types."""


def _fortran_e(x: float, width: int, digits: int) -> str:
    """Fortran Ew.d edit descriptor: [-]0.ddddE+xx right-justified."""
    x = float(x)
    if x == 0.0:
        m, exp = 0.0, 0
    else:
        exp = int(np.floor(np.log10(abs(x)))) + 1
        m = round(abs(x) / 10.0 ** exp, digits)
        if m >= 1.0:
            m /= 10.0
            exp += 1
    s = ("-" if x < 0 else "") + "0." + ("%.*f" % (digits, m))[2:] + "E%+03d" % exp
    return s.rjust(width)


def write_scfv(path, system, coeffs_up, coeffs_dn=None, rho=None):
    """Write coefficients (and optional rho columns [potDim, 3]) in the
    reference's scfV.dat layout."""
    coeffs_dn = coeffs_up if coeffs_dn is None else coeffs_dn
    alphas = np.concatenate([p.alphas for p in system.pot_types])
    if rho is None:
        rho = np.zeros((len(alphas), 3))
    lines = ["NUM_TYPES%5d" % len(system.pot_types)]
    for spin, coeffs in ((0, coeffs_up), (1, coeffs_dn)):
        lines.append("TOTAL__OR__SPIN_UP" if spin == 0 else "SPIN_DN")
        j = 0
        for p in system.pot_types:
            lines.append("%5d" % len(p.alphas))
            for _ in p.alphas:
                lines.append(" %s %s %s %s %s" % (
                    _fortran_e(coeffs[j], 17, 10), _fortran_e(alphas[j], 17, 10),
                    _fortran_e(rho[j, 0], 13, 6), _fortran_e(rho[j, 1], 13, 6),
                    _fortran_e(rho[j, 2], 13, 6)))
                j += 1
    lines.append("NUM__PLUSUJ__TERMS")
    lines.append("%5d" % 0)
    lines.append("TOTAL__OR__SPIN_UP")
    lines.append("SPIN_DN")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_scfv(path, system, spin: int = 1) -> np.ndarray:
    """Potential coefficients [potDim, spin] as initPotCoeffs reads them."""
    with open(path) as f:
        toks = [ln.split() for ln in f if ln.strip()]
    pos = 1                                   # NUM_TYPES header
    pot_dim = system.pot_dim
    out = np.zeros((pot_dim, spin))
    for s in range(spin):
        pos += 1                              # spin tag
        term = 0
        for ti, p in enumerate(system.pot_types):
            count = int(toks[pos][0])
            pos += 1
            if count != len(p.alphas):
                raise ScfvMismatch("scfV claims type %d has %d terms, the system has %d"
                                   % (ti + 1, count, len(p.alphas)))
            for _ in range(count):
                out[term, s] = float(toks[pos][0].replace("D", "E"))
                pos += 1
                term += 1
    return out
