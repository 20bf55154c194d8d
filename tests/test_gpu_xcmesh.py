"""imago_xc_mesh / imago_xc_rho_op / imago_xc_overlap (exchCorr.f90) on the
GPU against the oracle: ray counts exact, radial weights and every
operator value (LDA and GGA) at 1e-10 relative with the same zero
pattern, exchCorrOverlap at 1e-10 relative (normwise), on the small cell
(default 100-vector mesh), on the 64-atom Si supercell (8 vectors, so the
oracle's dense overlap stays within seconds) and on a-SiO2 sites."""
import numpy as np
import pytest

from conftest import rel_err
from imago_amd import configs
from imago_amd.system import XcMesh

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _check_sites(ctx, om, sites, gga=False):
    for site in sites:
        n, rw, op = ctx.xc_rho_op(site, gga=gga)
        nr, rwr, opr = om.site(site, gga=gga)
        assert n == nr == om.npts[site], (site, n, nr)
        assert rel_err(rw[:n], rwr[:n]) < 1e-14, site
        assert np.all(rw[n:] == 0.0)
        for o in range(op.shape[2]):
            assert rel_err(op[:, :, o], opr[:, :, o]) < TOL, (site, o)
            assert np.array_equal(op[:, :, o] == 0.0, opr[:, :, o] == 0.0), (site, o)


@pytest.mark.parametrize("two_elements", [True, False])
def test_small(gpu_ctx, oracle_mod, two_elements):
    s = configs.small_test_system(two_elements=two_elements)
    print("[gpu] xc mesh small two_elements=%s" % two_elements, flush=True)
    gpu_ctx.set_system(s)
    mx = gpu_ctx.xc_mesh()
    om = oracle_mod.XcOracleMesh(s)
    assert mx == om.maxpts
    _check_sites(gpu_ctx, om, range(len(s.pot_site_pos)))
    _check_sites(gpu_ctx, om, [0], gga=True)
    got = gpu_ctx.xc_overlap()
    ref = om.overlap()
    assert np.all(np.tril(got, -1) == 0.0)
    assert rel_err(got, ref) < TOL, rel_err(got, ref)


def test_si64_overlap(gpu_ctx, oracle_mod):
    s = configs.si64_supercell()
    mesh = XcMesh(num_sample_vectors=8)
    print("[gpu] xc mesh si64 (8 vectors)", flush=True)
    gpu_ctx.set_system(s)
    mx = gpu_ctx.xc_mesh(mesh)
    om = oracle_mod.XcOracleMesh(s, mesh)
    assert mx == om.maxpts
    _check_sites(gpu_ctx, om, (0, 37))
    got = gpu_ctx.xc_overlap()
    ref = om.overlap()
    assert rel_err(got, ref) < TOL, rel_err(got, ref)


def test_sio2_sites(gpu_ctx, oracle_mod):
    s = configs.sio2_amorphous()
    print("[gpu] xc mesh sio2 (100 vectors)", flush=True)
    gpu_ctx.set_system(s)
    mx = gpu_ctx.xc_mesh()
    om = oracle_mod.XcOracleMesh(s)
    assert mx == om.maxpts
    assert np.array_equal(np.array([gpu_ctx.xc_rho_op(i)[0] for i in (0, 100, 215)]),
                          om.npts[[0, 100, 215]])
    _check_sites(gpu_ctx, om, (0, 150))
    _check_sites(gpu_ctx, om, (80,), gga=True)


def test_errors(gpu_ctx):
    from imago_amd.lib import ImagoError
    s = configs.small_test_system()
    gpu_ctx.set_system(s)
    with pytest.raises(ImagoError):
        gpu_ctx.xc_rho_op(0)          # before the mesh
    with pytest.raises(ImagoError):
        gpu_ctx.xc_mesh(XcMesh(r_sample_space=1.5))
    gpu_ctx.xc_mesh(XcMesh(num_sample_vectors=4))
    with pytest.raises(ImagoError):
        gpu_ctx.xc_rho_op(99)


@pytest.mark.parametrize("name", ["small", "si64"])
def test_xc_potential(gpu_ctx, oracle_mod, name):
    if name == "small":
        s, mesh = configs.small_test_system(), XcMesh()
    else:
        s, mesh = configs.si64_supercell(), XcMesh(num_sample_vectors=8)
    print("[gpu] xc potential", name, flush=True)
    gpu_ctx.set_system(s)
    gpu_ctx.xc_mesh(mesh)
    om = oracle_mod.XcOracleMesh(s, mesh)
    rng = np.random.default_rng(9)
    rho = rng.normal(size=s.pot_dim) * 0.02 + 0.01
    core = np.abs(rng.normal(size=s.pot_dim)) * 0.005
    got = gpu_ctx.xc_potential(rho, core)
    ref = om.potential(rho, core)
    for k in range(3):
        assert rel_err(got[k], ref[k]) < TOL, (k, rel_err(got[k], ref[k]))
    got0 = gpu_ctx.xc_potential(rho)          # no core density: its row is 0
    assert np.all(got0[2] == 0.0)
    assert rel_err(got0[:2], got[:2]) < TOL


@pytest.mark.parametrize("name", ["small", "si64"])
def test_pot_solve(gpu_ctx, oracle_mod, name):
    """solveDPOSVX on the device against LAPACK dposvx('E', 'U'): the
    backward error at rounding level, and the solutions within the
    conditioning (the overlaps reach 1e10 on Si64)."""
    from scipy.linalg import lapack
    if name == "small":
        s, mesh = configs.small_test_system(), XcMesh()
    else:
        s, mesh = configs.si64_supercell(), XcMesh(num_sample_vectors=8)
    print("[gpu] pot solve", name, flush=True)
    gpu_ctx.set_system(s)
    es = gpu_ctx.elecstat()
    gpu_ctx.xc_mesh(mesh)
    xo = gpu_ctx.xc_overlap()
    rng = np.random.default_rng(3)
    for which, A in (("pot", es["potAlphaOverlap"]), ("xc", xo)):
        U = np.triu(A)
        Af = U + np.triu(A, 1).T
        b = rng.normal(size=(s.pot_dim, 2))
        x = gpu_ctx.pot_solve(which, b)
        res = np.abs(Af @ x - b).max() / (np.abs(Af).max() * np.abs(x).max())
        assert res < 1e-13, (which, res)
        out = lapack.dposvx(U, b, fact="E", lower=0)
        xr, rcond = out[5], out[6]
        cond = 1.0 / max(rcond, 1e-300)
        err = np.abs(x - xr).max() / np.abs(xr).max()
        assert err < max(1e-12, 1e-15 * cond), (which, err, cond)
        x1 = gpu_ctx.pot_solve(which, b[:, 0])   # the kept factor, one right-hand side
        assert np.allclose(x1, x[:, 0], rtol=1e-12, atol=1e-12 * np.abs(x).max())
