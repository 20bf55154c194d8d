"""Host-side pieces of the SCF loop, no GPU needed:

* imago_populate (populateStandard, populate.F90:190-485; host code in
  libimago_hip.so) against the oracle's restatement and against hand-made
  cases: filling order, a partly filled last level, degenerate levels at the
  Fermi level sharing the charge, the band gap and its kind;
* the iteration row of fort.7 (potentialUpdate.F90:1806-1808) and the
  driver's read-out of it (_harvest_result, src/scripts/imago.py:
  2630-2706), on the cases the reference's own tests use
  (src/tests/test_imago_api.py:113-240);
* the oracle's SCF on the small cell converges, with Tr(rho S) equal to
  the electron count in every iteration and the fitted charge holding
  exactly NUM_ELECTRONS after the constrained fit.
"""
import numpy as np
import pytest

from imago_amd import configs, lib, scf


def test_populate_matches_oracle_on_random_spectra(oracle_mod):
    rng = np.random.default_rng(3)
    for trial in range(50):
        nk = int(rng.integers(1, 6))
        ns = int(rng.integers(3, 12))
        ev = np.sort(rng.normal(size=(nk, ns)), axis=1)
        w = rng.uniform(0.5, 1.5, size=nk)
        w = 2.0 * w / w.sum()
        ne = float(rng.integers(1, 2 * ns))
        pop, info = lib.populate(ev, w, ne)
        ref = oracle_mod.populate(ev, w, ne)
        assert np.allclose(pop, ref, atol=1e-14)
        assert pop.sum() == pytest.approx(ne, abs=1e-10)


def test_populate_degenerate_fermi_level_shares_charge():
    # one k-point (weight 2): levels -2, -1, -1, -1, 0; 4 electrons fill -2
    # and leave 2 electrons for the three degenerate -1 levels
    ev = np.array([[-2.0, -1.0, -1.0 + 1e-10, -1.0 - 1e-10, 0.0]])
    pop, (e_occ, gap, kind) = lib.populate(ev, np.array([2.0]), 4.0)
    assert pop[0, 0] == pytest.approx(2.0)
    assert np.allclose(pop[0, 1:4], 2.0 / 3.0)
    assert pop[0, 4] == 0.0
    assert e_occ == pytest.approx(-1.0, abs=1e-9)
    assert gap == pytest.approx(1.0, abs=1e-9)
    assert kind == 1


def test_populate_gap_kind_and_metal():
    # two k-points of weight 1, 2 electrons: -1.0 (k 1) and -0.9 (k 2) filled;
    # the lowest empty level 0.2 is at k 2 as well: a direct gap
    ev = np.array([[-1.0, 0.5], [-0.9, 0.2]])
    pop, (e_occ, gap, kind) = lib.populate(ev, np.array([1.0, 1.0]), 2.0)
    assert np.allclose(pop, [[1.0, 0.0], [1.0, 0.0]])
    assert e_occ == pytest.approx(-0.9)
    assert gap == pytest.approx(1.1)
    assert kind == 1           # -0.9 (k 2) then 0.2 (k 2): direct
    ev = np.array([[-1.0, 0.5], [-0.9, 0.6]])
    _, (_, gap, kind) = lib.populate(ev, np.array([1.0, 1.0]), 3.0)
    # 3 electrons: -1.0, -0.9, then 0.5 at k 1; the next level 0.6 is at k 2
    assert (gap, kind) == (pytest.approx(0.1), 2)
    ev = np.array([[-1.0, 0.0005], [-0.9, 0.0]])
    _, (_, gap, kind) = lib.populate(ev, np.array([1.0, 1.0]), 3.0)
    assert (gap, kind) == (0.0, 0)   # a gap below 1e-3 Ha is a metal


def test_populate_refuses_too_many_electrons():
    with pytest.raises(lib.ImagoError):
        lib.populate(np.array([[-1.0, 0.0]]), np.array([2.0]), 5.0)


def _rec(it, delta, energy, gap=0.1, kind=1):
    return dict(iteration=it, occupied_energy=-0.3, electron_diff=1e-3, testable_delta=delta,
                total_energy=energy, band_gap=gap, gap_kind=kind)


def test_fort7_row_format():
    row = scf.fort7_row(_rec(12, 3.2e-5, -239.89915095, 0.0123, 2))
    assert row == ("   12   -0.30000000    0.00100000    0.00003200     -239.89915095"
                   "        0.00000000        0.01230000   2")


def test_harvest_converged_and_not(tmp_path):
    p = tmp_path / "fort.7"
    scf.write_iteration_file(p, [_rec(1, 0.01, -10.0), _rec(2, 0.00005, -10.5)])
    r = scf.harvest_iteration_file(p, 1.0e-4)
    assert r["converged"] and r["iterations"] == 2
    assert r["total_energy"] == pytest.approx(-10.5)
    assert r["magnetization"] == 0.0
    assert r["gap_ev"] == pytest.approx(0.1 * 27.21138386)
    assert r["gap_kind"] == "direct"
    scf.write_iteration_file(p, [_rec(1, 0.01, -10.0), _rec(2, 0.005, -10.5)])
    assert not scf.harvest_iteration_file(p, 1.0e-4)["converged"]


def test_harvest_short_rows_and_metal_code(tmp_path):
    # the reference's 5-column fixture: no magnetization or gap columns
    p = tmp_path / "gs_iter-fb.dat"
    p.write_text("# iter c2 c3 delta energy\n1 0.0 0.0 0.01 -10.0\n2 0.0 0.0 0.00005 -10.5\n")
    r = scf.harvest_iteration_file(p, 1.0e-4)
    assert r["converged"] and r["iterations"] == 2
    assert r["magnetization"] is None and r["gap_ev"] is None and r["gap_kind"] is None
    p.write_text("1 0.0 0.0 0.00001 -5.0 0.0 0.0 0\n\n  \n")
    r = scf.harvest_iteration_file(p, 1.0e-4)
    assert r["gap_kind"] == "none" and r["gap_ev"] == 0.0


def test_oracle_scf_small_cell_converges_with_exact_charge(oracle_mod):
    import scf as oscf
    s = configs.small_test_system(seed=5)
    recs, c = oscf.run_scf(s, configs.pot_coefficients(s), max_iter=40)
    ne = oscf.num_electrons(s)
    assert recs[-1]["testable_delta"] < 1e-4
    assert abs(recs[-1]["energy_change"]) < 1e-6
    intg = oscf.int_consts(s)
    for r in recs:
        assert r["charge_trace"] == pytest.approx(ne, rel=1e-10)
        # the constrained fit holds exactly NUM_ELECTRONS (potentialUpdate.F90:300-305)
        assert float(intg @ r["rho_vale"]) == pytest.approx(ne, rel=1e-12)
        assert r["total_energy"] < 0.0


def test_oracle_scf_xc_from_stored_operators_matches_c_routine(oracle_mod):
    """The oracle's makeSCFPot takes the exchange-correlation terms from the
    stored site operators (as the reference reads exchRhoOp back from its
    HDF5 file); they equal the C restatement's accumulation (orc_xc_potential,
    which re-evaluates the operator), and the numpy Wigner functional equals
    orc_xc_wigner."""
    import scf as oscf
    s = configs.small_test_system(seed=5)
    pot = oscf.ScfPotential(s)
    rng = np.random.default_rng(2)
    for _ in range(3):
        rt = rng.normal(size=s.pot_dim) * 0.05 + 0.02
        rc = np.abs(rng.normal(size=s.pot_dim)) * 0.01
        a = pot.xc_potential(rt, rc)
        b = pot.xc.potential(rt, rc).T
        assert np.allclose(a, b, rtol=1e-11, atol=1e-13 * np.max(np.abs(b)))
    for rho in (0.0, 1e-6, 0.3, 2.7, 40.0):
        p, e = oscf.wigner(np.array([rho]))
        pc, ec = oracle_mod.xc_wigner(rho)
        assert p[0] == pytest.approx(pc, rel=1e-14, abs=1e-300)
        assert e[0] == pytest.approx(ec, rel=1e-14, abs=1e-300)
