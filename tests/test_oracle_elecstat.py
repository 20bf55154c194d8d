"""Pins the oracle's electrostatics (oracle.c: orc_error_fn, orc_elecstat,
orc_residual_q; elecStat.f90 and mathSubs.f90 of the reference).

- error_fn bit for bit against the reference's own mathSubs.f90 compiled
  here (oracle/_ref), and within its stated eleven digits of libm's erf;
- every matrix term of an isolated site pair against the physics it
  stands for, by numerical quadrature (Gaussian-pair overlap, Coulomb
  energy of a Gaussian with a neutralised Gaussian, Gaussian in the field
  of a Gaussian-screened nucleus, the on-site limits);
- which alpha pairs the reference's L-shaped walk keeps, on masks worked
  out by hand for alphas chosen to stop the walk at known places;
- the lattice sums against an independent loop (brute-force nearest
  lattice point, no L-shaped walk);
- residualQ against an independent Ewald sum of the periodic point-charge
  potential with a different splitting parameter.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from imago_amd import configs
from imago_amd.system import PotType, System, make_lattice, reciprocal_vectors

SQPI = math.sqrt(math.pi)


def _pot_system(rv, pos, pots, cell_radius, thresh=None):
    """Sites at pos with one small atom type each; real-space cells out to
    cell_radius (bohr)."""
    rng = np.random.default_rng(11)
    at = configs.make_atom_type(rng, configs.SI, (0, 1), nalpha=4, amin=0.3, amax=30.0)
    pos = np.asarray(pos, dtype=np.float64)
    dims, sizes = make_lattice(np.asarray(rv, dtype=np.float64), cell_radius ** 2)
    kw = {} if thresh is None else dict(log_elec_thresh=thresh)
    return System(real_vectors=np.asarray(rv, dtype=np.float64), atom_types=[at],
                  atom_type_of=[0] * len(pos), atom_pos=pos, pot_types=pots,
                  pot_site_pos=pos.copy(), kpoints=np.zeros((1, 3)), kpoint_weights=[2.0],
                  log_basis_fn_thresh=20.0, gamma=True, cell_dims=dims, cell_sizes=sizes,
                  **kw)


def _pt(alphas, sites, z=4.0, nuc=5.0, element=configs.SI):
    return PotType(element=element, alphas=np.array(alphas, dtype=np.float64), nuc_charge=z,
                   nuc_alpha=nuc, sites=list(sites))


# ---------------------------------------------------------------- error_fn
def test_error_fn_against_reference_build(oracle_mod):
    if oracle_mod.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-8, 8, 4001), rng.uniform(-7, 7, 4000),
                         [0.46875, -0.46875, 4.0, -4.0, 6.0, -6.0, 0.0,
                          np.nextafter(0.46875, 1), np.nextafter(4.0, 5), np.nextafter(6.0, 7)]])
    for x in xs:
        assert oracle_mod.error_fn(x) == oracle_mod.error_fn(x, source="ref"), x


def test_error_fn_accuracy_and_branches(oracle_mod):
    xs = np.linspace(-7.5, 7.5, 30001)
    err = max(abs(oracle_mod.error_fn(x) - math.erf(x)) for x in xs)
    assert err < 5e-12, err
    # continuity at the range boundaries and oddness
    for b in (0.46875, 4.0, 6.0):
        lo, hi = oracle_mod.error_fn(np.nextafter(b, 0)), oracle_mod.error_fn(np.nextafter(b, 9))
        assert abs(hi - lo) < 1e-11, b
    for x in (0.1, 0.9, 3.0, 5.5, 6.5):
        assert oracle_mod.error_fn(-x) == -oracle_mod.error_fn(x)
    assert oracle_mod.error_fn(0.0) == 0.0


# ------------------------------------------------- isolated pair, by quadrature
def _radial_pair(b, r, V):
    """Integral of exp(-b |x - B|^2) V(|x - A|) over space, |A - B| = r, V
    spherical: (pi / (b r)) int_0^inf V(s) s (e^{-b(s-r)^2} - e^{-b(s+r)^2}) ds."""
    f = lambda s: V(s) * s * (math.exp(-b * (s - r) ** 2) - math.exp(-b * (s + r) ** 2))
    val, _ = integrate.quad(f, 0.0, r + 12.0 / math.sqrt(b), limit=400, epsabs=0, epsrel=1e-13)
    return math.pi / (b * r) * val


def _overlap_quad(a, b, A, B):
    out = 1.0
    for x in range(3):
        f = lambda t: math.exp(-a * (t - A[x]) ** 2 - b * (t - B[x]) ** 2)
        c = (a * A[x] + b * B[x]) / (a + b)
        v, _ = integrate.quad(f, c - 40.0 / math.sqrt(a + b), c + 40.0 / math.sqrt(a + b),
                              epsabs=0, epsrel=1e-13, limit=200)
        out *= v
    return out


def _neutral_coulomb(a, b, r):
    """Coulomb energy of exp(-a|x-A|^2) with exp(-b|x-B|^2) - (pi/b)^1.5
    delta_B: the potential of the first, (pi/a)^1.5 erf(sqrt(a) s)/s, on
    the neutral pair."""
    qa, qb = (math.pi / a) ** 1.5, (math.pi / b) ** 1.5
    Va = lambda s: qa * math.erf(math.sqrt(a) * s) / s if s > 0 else qa * 2 * math.sqrt(a) / SQPI
    return _radial_pair(b, r, Va) - qb * Va(r)


def _screened_nucleus_exact(a, c, r):
    """Integral of exp(-a|x-A|^2) exp(-c|x-B|^2) / |x-B|, |A - B| = r."""
    return _radial_pair(a, r, lambda s: math.exp(-c * s * s) / s)


CASES = [
    # alphas of site i's type, site j's type, hand-derived kept mask [b][a]
    # (row = alpha of site j, column = alpha of site i) at r = 2.5 bohr
    ("walk passes, tail below (n1 > n2)", [0.2, 1.0, 30.0], [0.3, 2.0], None),
    ("walk passes, tail right (n1 < n2)", [0.2, 1.0], [0.3, 2.0, 5.0], None),
    ("walk stops at l = 1", [0.2, 20.0, 200.0], [0.3, 30.0, 300.0],
     [[1, 1, 1], [1, 0, 0], [1, 0, 0]]),
    ("tail kept past the threshold", [0.2, 50.0], [0.3, 60.0], None),
]


@pytest.mark.parametrize("name,al1,al2,mask", CASES, ids=[c[0] for c in CASES])
def test_isolated_pair_terms(oracle_mod, name, al1, al2, mask):
    r = 2.5
    A = np.array([10.0, 11.0, 12.0])
    d = np.array([1.0, 2.0, 2.0]) / 3.0
    B = A + r * d
    z2, nuc2 = 6.0, 7.0
    s = _pot_system(60.0 * np.eye(3), [A, B],
                    [_pt(al1, [0]), _pt(al2, [1], z=z2, nuc=nuc2, element=configs.O)], 70.0)
    res = oracle_mod.elecstat(s)
    n1, n2 = len(al1), len(al2)
    if mask is None:
        mask = np.ones((n2, n1), dtype=int)
    mask = np.asarray(mask)
    ovl, nl, lo = res["potAlphaOverlap"], res["nonLocalNeutQPot"], res["localNeutQPot"]
    for ia, a in enumerate(al1):
        for ib, b in enumerate(al2):
            row, col = n1 + ib, ia          # site j's term, site i's term
            if not mask[ib, ia]:
                assert ovl[row, col] == 0.0 and nl[row, col] == 0.0, (ia, ib)
                continue
            ref_ov = _overlap_quad(a, b, A, B)
            assert ovl[row, col] == pytest.approx(ref_ov, rel=1e-10, abs=1e-300), (ia, ib)
            ref_nl = _neutral_coulomb(a, b, r)
            # error_fn carries ~2.4e-12 absolute error into the difference
            # of two erf values times (pi/a)^1.5 (pi/b)^1.5 / r
            scale = (math.pi / a) ** 1.5 * (math.pi / b) ** 1.5 / r
            assert abs(nl[row, col] - ref_nl) <= 1e-11 * scale + 1e-9 * abs(ref_nl), (
                ia, ib, nl[row, col], ref_nl)
            assert lo[row, col] == 0.0
    # on-site (i = j, cell 0) limits: Gaussian-Gaussian Coulomb energy at
    # zero separation, sum over every alpha pair of the type with itself
    for ia, a in enumerate(al1):
        for ib, b in enumerate(al1):
            ref = (math.pi / a) ** 1.5 * (math.pi / b) ** 1.5 * 2 * math.sqrt(a * b / (a + b)) / SQPI
            q, _ = integrate.quad(lambda x: 4 * math.pi * x * x * math.exp(-b * x * x) *
                                  (math.pi / a) ** 1.5 * math.erf(math.sqrt(a) * x) / x,
                                  0, 40 / math.sqrt(b), epsabs=0, epsrel=1e-13, limit=200)
            assert ref == pytest.approx(q, rel=1e-10)
            assert lo[ib, ia] == pytest.approx(ref, rel=1e-12)
    # nuclear: site i's alphas in the field of both screened nuclei; the
    # on-site nucleus is the exact limit 2 pi / (a + c)
    z1, nuc1 = 4.0, 5.0
    nln, lon = res["nonLocalNucQPot"], res["localNucQPot"]
    th = s.log_elec_thresh
    for ia, a in enumerate(al1):
        want = -z1 * 2 * math.pi / (a + nuc1)
        if r * r * a * nuc2 / (a + nuc2) < th:
            want += -z2 * _screened_nucleus_exact(a, nuc2, r)
        assert nln[ia] == pytest.approx(want, rel=1e-9), (ia, nln[ia], want)
        q, _ = integrate.quad(lambda x: 4 * math.pi * x * math.exp(-a * x * x), 0, 60 / math.sqrt(a),
                              epsabs=0, epsrel=1e-13)
        assert lon[ia] == pytest.approx(-z1 * q, rel=1e-12)


def test_nuclear_walk_stops_at_first_far_alpha(oracle_mod):
    # screens r^2 a c / (a + c) of 0.1, 10, 40 (> th = 36.8), 2 at r = 1:
    # the walk stops at the third alpha, so the fourth is not added even
    # though its own screen is small
    c = 50.0
    scr = [0.1, 10.0, 40.0, 2.0]
    al = [x * c / (c - x) for x in scr]
    order = np.argsort(al)
    assert list(order) != list(range(4))   # not ascending on purpose
    A = np.array([10.0, 10.0, 10.0])
    s = _pot_system(60.0 * np.eye(3), [A, A + np.array([1.0, 0, 0])],
                    [_pt(al, [0], z=0.0), _pt([1.0], [1], z=3.0, nuc=c, element=configs.O)],
                    70.0)
    res = oracle_mod.elecstat(s)
    nln = res["nonLocalNucQPot"][:4]
    for m in range(4):
        want = -3.0 * _screened_nucleus_exact(al[m], c, 1.0) if m < 2 else 0.0
        assert nln[m] == pytest.approx(want, rel=1e-9, abs=0), m


# -------------------------------------------- lattice sums, independent loop
def _nearest_lattice_point(rv, v):
    best, bl = None, None
    f = np.linalg.solve(rv, v)
    base = np.round(f)
    for n in np.ndindex(5, 5, 5):
        L = rv @ (base + np.array(n) - 2)
        d = np.sum((v - L) ** 2)
        if best is None or d < best - 1e-12:
            best, bl = d, L
    return bl


def _brute_elecstat(s):
    th = s.log_elec_thresh
    terms = s.pot_terms()
    P = len(terms)
    col0 = {}
    for q, (t, a) in enumerate(terms):
        col0.setdefault(t, q)
    ovl, nl, lo = np.zeros((P, P)), np.zeros((P, P)), np.zeros((P, P))
    site_type = {i: t for t, p in enumerate(s.pot_types) for i in p.sites}
    for i in range(len(s.pot_site_pos)):
        t1 = site_type[i]
        a1 = s.pot_types[t1].alphas
        for j in range(len(s.pot_site_pos)):
            t2 = site_type[j]
            a2 = s.pot_types[t2].alphas
            sep = s.pot_site_pos[i] - s.pot_site_pos[j]
            ms = sep - _nearest_lattice_point(s.real_vectors, sep)
            minn = min(len(a1), len(a2))
            for R in s.cell_dims:
                r2 = float(np.sum((ms - R) ** 2))
                r = math.sqrt(r2)
                diag = [a1[l] * a2[l] / (a1[l] + a2[l]) * r2 for l in range(minn)]
                for a in range(len(a1)):
                    for b in range(len(a2)):
                        q = min(a, b, minn - 2)
                        if any(diag[l] > th for l in range(q + 1)):
                            continue
                        x, y = a1[a], a2[b]
                        red = x * y / (x + y)
                        cc = math.pi ** 3 / (x * y) ** 1.5
                        at = (col0[t2] + b, col0[t1] + a)
                        if r2 >= 1e-8:
                            nl[at] += cc / r * (math.erf(math.sqrt(red) * r) - math.erf(math.sqrt(x) * r))
                        else:
                            lo[at] += cc * 2 / SQPI * math.sqrt(red)
                        ovl[at] += (math.pi / (x + y)) ** 1.5 * math.exp(-red * r2)
    return ovl, nl, lo


@pytest.mark.parametrize("two_elements", [True, False])
def test_lattice_sums_independent_loop(oracle_mod, two_elements):
    s = configs.small_test_system(two_elements=two_elements)
    res = oracle_mod.elecstat(s)
    ovl, nl, lo = _brute_elecstat(s)
    # libm erf here, the reference's eleven-digit error_fn in the oracle
    assert np.max(np.abs(res["potAlphaOverlap"] - ovl)) <= 1e-13 * np.max(np.abs(ovl))
    assert np.max(np.abs(res["nonLocalNeutQPot"] - nl)) <= 1e-9 * np.max(np.abs(nl))
    assert np.max(np.abs(res["localNeutQPot"] - lo)) <= 1e-13 * np.max(np.abs(lo))


def test_overlap_symmetric_when_the_cells_cover_the_range(oracle_mod):
    s = configs.small_test_system()
    dims, sizes = make_lattice(s.real_vectors, 30.0 ** 2)
    s.cell_dims, s.cell_sizes = dims, sizes
    ov = oracle_mod.elecstat(s)["potAlphaOverlap"]
    assert np.max(np.abs(ov - ov.T)) <= 1e-13 * np.max(np.abs(ov))
    # Gram matrix of the Gamma-point Bloch sums of the Gaussians
    w = np.linalg.eigvalsh(0.5 * (ov + ov.T))
    assert w[0] > -1e-12 * w[-1]


# ------------------------------------------------------- residualQ, by Ewald
def _ewald_phi(rv, s, eta, nmax=3, gmax=9):
    """Periodic potential of unit point charges at the lattice points with
    a neutralising background, at s (not a lattice point), parameter eta."""
    recip, vol = reciprocal_vectors(rv)
    real = 0.0
    for n in np.ndindex(2 * nmax + 1, 2 * nmax + 1, 2 * nmax + 1):
        R = rv @ (np.array(n) - nmax)
        d = np.linalg.norm(s - R)
        real += special.erfc(math.sqrt(eta) * d) / d
    rec = 0.0
    for n in np.ndindex(2 * gmax + 1, 2 * gmax + 1, 2 * gmax + 1):
        if n == (gmax, gmax, gmax):
            continue
        G = recip @ (np.array(n) - gmax)
        g2 = G @ G
        rec += math.exp(-g2 / (4 * eta)) / g2 * math.cos(G @ s)
    return real + 4 * math.pi / vol * rec - math.pi / (vol * eta)


def _ewald_phi_reg0(rv, eta, nmax=3, gmax=9):
    """lim_{s->0} phi(s) - 1/s."""
    recip, vol = reciprocal_vectors(rv)
    real = 0.0
    for n in np.ndindex(2 * nmax + 1, 2 * nmax + 1, 2 * nmax + 1):
        if n == (nmax, nmax, nmax):
            continue
        d = np.linalg.norm(rv @ (np.array(n) - nmax))
        real += special.erfc(math.sqrt(eta) * d) / d
    rec = 0.0
    for n in np.ndindex(2 * gmax + 1, 2 * gmax + 1, 2 * gmax + 1):
        if n == (gmax, gmax, gmax):
            continue
        G = recip @ (np.array(n) - gmax)
        rec += math.exp(-(G @ G) / (4 * eta)) / (G @ G)
    return real + 4 * math.pi / vol * rec - math.pi / (vol * eta) - 2 * math.sqrt(eta / math.pi)


def test_residual_q_equals_ewald_potential(oracle_mod):
    """residualQ's element is, for each site pair, af * (phi(s) - sum_R
    erfc(sqrt(red) r)/r + pi / (V a)) with phi the periodic point-charge
    potential: the real- and reciprocal-space parts at the smallest alpha
    combine into phi whatever the splitting parameter.  The reference's
    real-space loop stops at the first cell (in |R| order) beyond both
    ranges (elecStat.f90:963-968), which can come before closer images of
    a pair whose minimal separation is long; those skipped terms are taken
    off here, found with a brute-force nearest lattice point."""
    rv = np.array([[7.0, 0.0, 0.0], [0.8, 6.5, 0.0], [0.3, 0.6, 7.6]]).T
    pos = [[0.4, 0.3, 0.2], [3.1, 3.3, 3.9]]
    al1, al2 = [0.25, 1.5, 9.0], [0.35, 4.0]
    s = _pot_system(rv, pos, [_pt(al1, [0]), _pt(al2, [1], element=configs.O)], 28.0)
    got = oracle_mod.residual_q(s)
    _, vol = reciprocal_vectors(rv)
    pi32 = math.pi ** 1.5
    sqth = math.sqrt(s.log_elec_thresh)
    min_pot = 0.25
    alphas = [al1, al2]
    pos = np.array(pos)
    skipped_any = False
    for row, t2 in enumerate((0, 1)):
        for col, (t1, k) in enumerate(s.pot_terms()):
            a, b0 = alphas[t1][k], alphas[t2][0]
            af = pi32 / a ** 1.5 * pi32 / b0 ** 1.5
            red = a * b0 / (a + b0)
            mra = a * min_pot / (a + min_pot)
            sv = pos[t1] - pos[t2]
            erfc_sum = 0.0
            for n in np.ndindex(9, 9, 9):
                R = rv @ (np.array(n) - 4)
                d = np.linalg.norm(sv - R)
                if d > 1e-9:
                    erfc_sum += special.erfc(math.sqrt(red) * d) / d
            if t1 == t2:
                phi = _ewald_phi_reg0(rv, 0.3) + 2 * math.sqrt(red / math.pi)
            else:
                phi = _ewald_phi(rv, sv, 0.3)
            ms = sv - _nearest_lattice_point(rv, sv)
            skipped, stopped = 0.0, False
            for R in s.cell_dims:
                d = float(np.linalg.norm(ms - R))
                if d < 1e-4:
                    continue
                stopped = stopped or (math.sqrt(mra) * d > sqth and math.sqrt(red) * d > sqth)
                if stopped:
                    skipped += (math.erf(math.sqrt(red) * d) - math.erf(math.sqrt(mra) * d)) / d
            skipped_any = skipped_any or abs(skipped) > 1e-6
            want = af * (phi - erfc_sum + math.pi / (vol * a) - skipped)
            assert got[row, col] == pytest.approx(want, rel=2e-9), (row, col, got[row, col], want)
    assert skipped_any   # the early stop is exercised
