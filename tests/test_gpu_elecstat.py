"""imago_elecstat / imago_residual_q (elecStat.f90) on the GPU against the
oracle: every output at 1e-10 relative (normwise per array), the same
zero pattern where the reference's walks stop, on small systems with
unequal alpha counts and multi-site types, on the 64-atom Si supercell in
full, and on the 216-atom a-SiO2 cell (residualQ on a sample of sites: the
oracle needs about 40 s for all of them)."""
import numpy as np
import pytest

from conftest import rel_err
from imago_amd import configs
from imago_amd.system import PotType, System, make_lattice

pytestmark = pytest.mark.gpu

TOL = 1e-10
KEYS = ("potAlphaOverlap", "nonLocalNeutQPot", "localNeutQPot", "nonLocalNucQPot",
        "localNucQPot")


def _check_elecstat(ctx, oracle_mod, s, label):
    ctx.set_system(s)
    got = ctx.elecstat()
    ref = oracle_mod.elecstat(s)
    for k in KEYS:
        assert got[k].shape == ref[k].shape, (label, k)
        assert rel_err(got[k], ref[k]) < TOL, (label, k, rel_err(got[k], ref[k]))
        # the same entries are left out by the screening walks
        assert np.array_equal(got[k] == 0.0, ref[k] == 0.0), (label, k)
    return got


def _check_residual(ctx, oracle_mod, s, label, sites=None):
    got = ctx.residual_q()
    ref = oracle_mod.residual_q(s, sites=sites)
    if sites is not None:
        cols = [q for q, (t, a) in enumerate(s.pot_terms())
                if any(sites[0] <= i < sites[1] for i in s.pot_types[t].sites)]
        got, ref = got[:, cols], ref[:, cols]
    assert rel_err(got, ref) < TOL, (label, rel_err(got, ref))


@pytest.mark.parametrize("two_elements", [True, False])
def test_small_systems(gpu_ctx, oracle_mod, two_elements):
    s = configs.small_test_system(two_elements=two_elements)
    print("[gpu] elecstat small two_elements=%s potDim=%d" % (two_elements, s.pot_dim),
          flush=True)
    _check_elecstat(gpu_ctx, oracle_mod, s, "small")
    _check_residual(gpu_ctx, oracle_mod, s, "small")


def test_walk_stops_and_wide_types(gpu_ctx, oracle_mod):
    """An isolated pair whose L-shaped walk stops at l = 1, a type with 40
    alphas (two reciprocal-table widths), and a long cell list."""
    rng = np.random.default_rng(11)
    at = configs.make_atom_type(rng, configs.SI, (0, 1), nalpha=4, amin=0.3, amax=30.0)
    rv = np.array([[9.0, 0.2, 0.0], [0.0, 8.5, 0.3], [0.4, 0.0, 9.5]]).T
    pos = np.array([[0.5, 0.4, 0.3], [2.0, 1.9, 2.4], [5.1, 4.2, 6.0]])
    pots = [PotType(element=configs.SI, alphas=np.array([0.2, 20.0, 200.0]), nuc_charge=4.0,
                    nuc_alpha=5.0, sites=[0]),
            PotType(element=configs.O, alphas=np.geomspace(0.3, 3.0e3, 40), nuc_charge=6.0,
                    nuc_alpha=7.0, sites=[1, 2])]
    dims, sizes = make_lattice(rv, 26.0 ** 2)
    s = System(real_vectors=rv, atom_types=[at], atom_type_of=[0, 0, 0], atom_pos=pos,
               pot_types=pots, pot_site_pos=pos.copy(), kpoints=np.zeros((1, 3)),
               kpoint_weights=[2.0], log_basis_fn_thresh=20.0, gamma=True, cell_dims=dims,
               cell_sizes=sizes)
    print("[gpu] elecstat wide types potDim=%d" % s.pot_dim, flush=True)
    got = _check_elecstat(gpu_ctx, oracle_mod, s, "wide")
    assert np.count_nonzero(got["potAlphaOverlap"] == 0.0) > 0   # some pairs screened out
    _check_residual(gpu_ctx, oracle_mod, s, "wide")


def test_si64(gpu_ctx, oracle_mod):
    s = configs.si64_supercell()
    print("[gpu] elecstat si64 potDim=%d" % s.pot_dim, flush=True)
    _check_elecstat(gpu_ctx, oracle_mod, s, "si64")
    print("[gpu] residualQ si64", flush=True)
    _check_residual(gpu_ctx, oracle_mod, s, "si64")


def test_sio2(gpu_ctx, oracle_mod):
    s = configs.sio2_amorphous()
    print("[gpu] elecstat sio2 potDim=%d" % s.pot_dim, flush=True)
    _check_elecstat(gpu_ctx, oracle_mod, s, "sio2")
    for sites in ((0, 3), (70, 74), (213, 216)):   # Si, the Si/O boundary, O
        print("[gpu] residualQ sio2 sites", sites, flush=True)
        _check_residual(gpu_ctx, oracle_mod, s, "sio2", sites=sites)


def test_errors(gpu_ctx):
    from imago_amd.lib import ImagoError
    s = configs.small_test_system()
    gpu_ctx.set_system(s)
    with pytest.raises(ImagoError):
        gpu_ctx.elecstat(log_elec_thresh=-1.0)
    dims, sizes = s.recip_cells()
    bad = sizes.copy()
    bad[1] = 0.0
    with pytest.raises(ImagoError):
        gpu_ctx.residual_q(recip_cells=(dims, bad))
