"""Post-SCF density of states and band structure on the GPU.

imago_dos (Mulliken populations, channel sums and Gaussian broadening on
the device) against the oracle's computeDOS restatement (oracle/dos.py) on
the same eigenpairs; the whole density_of_states driver on the Si2
primitive cell with its 10 k-points through the direct EP path; and the
band structure, whose eigenvalues at the high-symmetry points equal those
of a separate solve at those k-points alone."""
import numpy as np
import pytest

import dos as odos
from imago_amd import configs, lib, pscf, scf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gamma", [False, True])
@pytest.mark.parametrize("detail", [0, 1, 3])
def test_dos_matches_oracle(gpu_ctx, oracle_mod, gamma, detail):
    s = configs.small_test_system(seed=9, gamma=gamma)
    print("[gpu] dos small gamma=%s detail=%d" % (gamma, detail), flush=True)
    gpu_ctx.set_system(s)
    for kind in ("ol", "ke", "np", "ep"):
        gpu_ctx.assemble(kind)
    n = s.vale_dim
    coeffs = configs.pot_coefficients(s, seed=3)
    ev = gpu_ctx.secular(coeffs, n)
    occ, (e_occ, _, _) = lib.populate(ev, s.kpoint_weights, 6.0)
    E = ev - e_occ
    chan, nch = pscf.pdos_channels(s, detail)
    emin, emax, delta, sigma = -2.0, 2.0, 0.004, 0.02
    npts = int((emax - emin) / delta)
    pd, en, loc = gpu_ctx.dos(E, 1, chan, nch, emin, delta, npts, sigma)
    S = oracle_mod.assemble(s, "ol")
    C = [gpu_ctx.eigenvectors(k, n) for k in range(s.num_kpoints)]
    Sf = [oracle_mod.unpack(S[k], n) for k in range(s.num_kpoints)]
    ls = [[l for l, _ in t.shells] for t in s.atom_types]
    idx, tot = odos.pdos_index(list(s.atom_type_of), ls, [t.num_states for t in s.atom_types],
                               detail)
    assert tot == nch and np.array_equal(idx, chan)
    ref = odos.compute_dos(E, C, Sf, s.kpoint_weights, 1, idx, tot, emin, emax, delta, sigma,
                           float(occ.sum()))
    scale = np.abs(ref["pdos_raw"]).max()
    assert np.abs(pd - ref["pdos_raw"]).max() <= 1e-10 * scale
    assert np.allclose(en, ref["electron_number_raw"], rtol=1e-10, atol=1e-12)
    assert np.allclose(loc, ref["localization"], rtol=1e-10, atol=1e-12)


def test_density_of_states_si2_ten_kpoints(gpu_ctx):
    s = configs.si2_primitive()
    coeffs = configs.pot_coefficients(s, seed=3)
    print("[gpu] dos si2 %d k-points" % s.num_kpoints, flush=True)
    res = pscf.density_of_states(gpu_ctx, s, coeffs, sigma_ev=0.1, emin_ev=-60.0,
                                 emax_ev=60.0, delta_ev=0.02, detail=2)
    assert np.allclose(res.pdos.sum(axis=0), res.tdos, rtol=1e-12, atol=1e-12)
    assert res.electrons_expected == pytest.approx(scf.num_electrons(s), rel=1e-12)
    assert 0.5 < res.electron_factor < 2.0
    # every state of the window counted once by the integrated TDOS, up to
    # the broadening tails at its ends
    assert res.integrated_area / res.electron_factor == pytest.approx(res.states_in_range,
                                                                        rel=0.02)
    assert np.all(res.localization > 0)


def test_band_structure_matches_separate_solves(gpu_ctx):
    s = configs.si2_primitive()
    coeffs = configs.pot_coefficients(s, seed=3)
    G, X, L = [0, 0, 0], [0.5, 0, 0.5], [0.5, 0.5, 0.5]
    print("[gpu] bands si2", flush=True)
    res = pscf.band_structure(gpu_ctx, s, coeffs, [[L, G, X]], 24)
    assert res.eigenvalues.shape == (len(res.kpoints), s.vale_dim)
    hs = res.kpoints[res.high_symmetry]
    s1 = pscf.with_kpoints(s, hs, np.full(len(hs), 2.0 / len(hs)))
    gpu_ctx.set_system(s1)
    for kind in ("ol", "ke", "np", "ep_direct"):
        gpu_ctx.assemble(kind)
    ev = gpu_ctx.secular(coeffs, s.vale_dim)
    got = res.eigenvalues[res.high_symmetry] + res.occupied_energy
    assert np.max(np.abs(got - ev)) < 1e-9
