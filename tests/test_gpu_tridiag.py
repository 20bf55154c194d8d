"""The batched tridiagonal divide and conquer on the GPU (imago_tridiag_eig,
the default tridiagonal stage of imago_secular's eigensolve): every case of
tests/tridiag_cases.py against LAPACK (eigenvalues, orthogonality and
residual at 1e-12 of max |T|), batches of different matrices of one order,
the orders of the workloads (896 and 2048), and the eigenvectors of well
separated eigenvalues against LAPACK's up to sign."""
import numpy as np
import pytest
import scipy.linalg as sl

import tridiag_cases

pytestmark = pytest.mark.gpu


def test_cases_match_lapack(gpu_ctx):
    for name, d, e in tridiag_cases.cases(max_n=300):
        print("[gpu] tridiag %s n=%d" % (name, len(d)), flush=True)
        w, Z = gpu_ctx.tridiag_eig(d, e)
        ev, orth, res = tridiag_cases.check(d, e, w[0], Z[0])
        assert np.all(np.diff(w[0]) >= 0), name
        assert ev < 1e-12 and orth < 1e-12 and res < 1e-12, (name, ev, orth, res)


@pytest.mark.parametrize("n,nb", [(64, 3), (127, 4), (896, 4), (2048, 1)])
def test_batches_and_workload_orders(gpu_ctx, n, nb):
    rng = np.random.default_rng(n + nb)
    d = rng.normal(size=(nb, n)) * np.linspace(1.0, 50.0, nb)[:, None]
    e = rng.normal(size=(nb, n - 1))
    print("[gpu] tridiag batch n=%d nb=%d" % (n, nb), flush=True)
    w, Z = gpu_ctx.tridiag_eig(d, e)
    print("[gpu]   device time %.3f ms" % gpu_ctx.timing(17), flush=True)
    for b in range(nb):
        ev, orth, res = tridiag_cases.check(d[b], e[b], w[b], Z[b])
        assert ev < 1e-12 and orth < 1e-12 and res < 1e-12, (b, ev, orth, res)
    # eigenvectors of separated eigenvalues agree with LAPACK's up to sign
    wr, Zr = sl.eigh_tridiagonal(d[0], e[0])
    gap = np.minimum(np.r_[np.inf, np.diff(wr)], np.r_[np.diff(wr), np.inf])
    sep = gap > 1e-3 * np.abs(d[0]).max()
    dots = np.abs(np.sum(Z[0][:, sep] * Zr[:, sep], axis=0))
    assert np.all(np.abs(dots - 1.0) < 1e-9)


def test_graph_replay_matches_direct_launches(gpu_ctx):
    """The divide and conquer replayed from its captured graph gives the
    same bits as its launches issued one by one: the first call captures,
    later calls replay with other work on the stream between them (the
    traces), and a system of another order captures again."""
    from imago_amd import configs
    for s in (configs.small_test_system(seed=4), configs.small_test_system(seed=5, gamma=True),
              configs.si64_supercell(reps=(1, 1, 1), mesh=(2, 2, 2), cutoff_bohr=16.0)):
        print("[gpu] graph replay %s n=%d" % (s.name, s.vale_dim), flush=True)
        gpu_ctx.set_system(s)
        for kind in ("ol", "ke", "np", "ep_direct"):
            gpu_ctx.assemble(kind)
        n = s.vale_dim
        coeffs = configs.pot_coefficients(s, seed=2)
        runs = []
        for graph in (1, 1, 1, 0):
            gpu_ctx.tune(graph=graph)
            ev = gpu_ctx.secular(coeffs, n)
            vec = [gpu_ctx.eigenvectors(k, n) for k in range(s.num_kpoints)]
            occ = np.zeros((s.num_kpoints, n))
            occ[:, :2] = 1.0
            gpu_ctx.density_traces(occ)
            runs.append((ev, vec))
        for ev, vec in runs[1:]:
            assert np.array_equal(ev, runs[0][0])
            for a, b in zip(vec, runs[0][1]):
                assert np.array_equal(a, b)
    gpu_ctx.tune(graph=1)
