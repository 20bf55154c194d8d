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


SCF_DRIVER = os.path.join(ROOT, "imago_amd", "fortran", "_bin", "scf_driver")


def _write_input(path, s, coeffs, max_iter=40, energy_tol=1e-6):
    """The module state of a Fortran host (types, sites, lattice, k-points,
    the MAIN_INPUT_DATA block, the XC mesh input, scfV.dat's coefficients)
    as the stream file scf_driver reads."""
    import numpy as np
    from imago_amd.system import XcMesh
    d = s.desc_arrays()
    dims, sizes = s.recip_cells()
    mesh = XcMesh()
    nel = sum(p.nuc_charge * len(p.sites) for p in s.pot_types)
    hdr = np.array([len(s.atom_types), d["max_alphas"], d["max_states"], s.num_atoms,
                    len(s.pot_types), d["max_pot_alphas"], len(s.pot_site_pos),
                    len(s.cell_sizes), s.num_kpoints, int(s.gamma), len(sizes), max_iter,
                    mesh.num_sample_vectors, 100, 2, 0], dtype=np.int32)
    dhdr = np.concatenate([[s.log_basis_fn_thresh, s.log_elec_thresh, nel, 0.2, 1e-4, energy_tol,
                            mesh.weight_in, mesh.weight_out, mesh.r_sample_in,
                            mesh.r_sample_out, mesh.r_sample_space], d["real_vectors"]])
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).ravel()
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
    parts = [hdr, f64(dhdr),
             i32(d["type_element"]), i32(d["type_num_orb_alphas"]), f64(d["type_alphas"]),
             i32(d["type_num_states"]), f64(d["type_basis"]), i32(d["type_lm_index"]),
             i32(d["atom_type"]), f64(d["atom_pos"]),
             i32(d["pot_type_element"]), i32(d["pot_type_num_alphas"]),
             f64(d["pot_type_alphas"]), f64(d["pot_type_nuc_charge"]),
             f64(d["pot_type_nuc_alpha"]), i32(d["pot_type_first_site"]),
             i32(d["pot_type_multiplicity"]),
             i32(d["pot_site_type"]), f64(d["pot_site_pos"]), f64(d["cell_dims"]),
             f64(d["cell_sizes"]), f64(d["kpoints"]), f64(d["kpoint_weights"]),
             f64(dims), f64(sizes), f64([p.covalent_radius for p in s.pot_types]),
             f64(coeffs)]
    with open(path, "wb") as f:
        for a in parts:
            f.write(a.tobytes())


@pytest.mark.parametrize("gamma", [False, True])
def test_fortran_host_runs_integrals_and_scf(tmp_path, oracle_mod, gamma):
    """A Fortran program fills imago_system_desc from its own arrays with
    c_loc, runs set-up (imago_set_system, OL/KE/NP/EP assembly,
    imago_get_packed) and the mainSCF loop (imago_secular, imago_populate,
    imago_density_traces, imago_scf_update) through O_ImagoHIP, writing
    fort.7 and fort.14 as the reference does.  The packed matrices match the
    oracle at 1e-10, the first iteration's eigenvalues LAPACK at 1e-8 Ha,
    and every iteration's total energy the oracle's SCF within the 1e-8 Ha
    of fort.7's format; the driver's read-out of fort.7 says converged."""
    import numpy as np
    import scf as oscf
    from conftest import rel_err
    from imago_amd import configs
    from imago_amd import scf as pscf
    if not os.path.exists(SCF_DRIVER):
        pytest.fail("scf_driver not built (imago_amd.build builds it)")
    s = configs.small_test_system(seed=5, gamma=gamma)
    coeffs = configs.pot_coefficients(s)
    inp = tmp_path / "system.bin"
    _write_input(inp, s, coeffs)
    print("[gpu] fortran scf driver gamma=%s" % gamma, flush=True)
    r = subprocess.run([SCF_DRIVER, str(inp)], capture_output=True, text=True, timeout=300,
                       cwd=tmp_path)
    unit20 = tmp_path / "fort.20"
    note = unit20.read_text() if unit20.exists() else ""
    # the binding reports a failure on unit 20 and stops (rc 0, "STOP" on stderr)
    assert r.returncode == 0 and not note.strip() and "STOP" not in r.stderr, \
        r.stdout + r.stderr + note
    n = s.vale_dim
    npk = n * (n + 1) // 2 * (1 if gamma else 2)
    raw = np.fromfile(tmp_path / "matrices.bin", dtype=np.float64)
    mats = raw[:4 * npk].reshape(4, npk)
    ev = raw[4 * npk:].reshape(s.num_kpoints, n)
    for q, kind in enumerate(("ol", "ke", "np")):
        ref = oracle_mod.assemble(s, kind, raw=True)[0]
        assert rel_err(mats[q], ref) < 1e-10, kind
    ref = oracle_mod.assemble(s, "ep", 0, 1, raw=True)[0][0]
    assert rel_err(mats[3], ref) < 1e-10
    S = oracle_mod.assemble(s, "ol")
    H = oracle_mod.hamiltonian(s, coeffs)
    for k in range(s.num_kpoints):
        w, _, info = oracle_mod.hegv(oracle_mod.unpack(H[k], n), oracle_mod.unpack(S[k], n))
        assert np.max(np.abs(ev[k] - w)) < 1e-8
    recs, _ = oscf.run_scf(s, coeffs, max_iter=40)
    rows = [l.split() for l in (tmp_path / "fort.7").read_text().splitlines()[1:] if l.strip()]
    # the stop test |dE| < 1e-6 can fall either side of the threshold by
    # rounding on the last step (the host here stores the EP panels, the
    # oracle sums them in another order): one iteration apart at most
    assert abs(len(rows) - len(recs)) <= 1
    for row, rec in zip(rows, recs):
        assert int(row[0]) == rec["iteration"]
        assert abs(float(row[4]) - rec["total_energy"]) < 1e-7
    out = pscf.harvest_iteration_file(tmp_path / "fort.7", 1.0e-4)
    assert out["converged"] and out["iterations"] == int(rows[-1][0])
    assert out["total_energy"] == pytest.approx(recs[-1]["total_energy"], abs=2e-6)
    e14 = [l.split() for l in (tmp_path / "fort.14").read_text().splitlines()[1:] if l.strip()]
    assert len(e14) == len(rows)
    assert float(e14[-1][4]) == pytest.approx(recs[-1]["total_energy"], rel=1e-6)
