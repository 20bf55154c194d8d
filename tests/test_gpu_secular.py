"""H(k) build, batched generalized eigensolve and density traces on the GPU
against the oracle (secularEqn.F90 / readPackedMatrixAccum restatement,
LAPACK zhegv/dsygv via scipy, matrixElementMult)."""
import numpy as np
import pytest

from conftest import rel_err
from imago_amd import configs
from imago_amd.lib import ImagoError

pytestmark = pytest.mark.gpu


def _setup(gpu_ctx, s):
    gpu_ctx.set_system(s)
    for kind in ("ol", "ke", "np", "ep"):
        gpu_ctx.assemble(kind)


@pytest.mark.parametrize("gamma", [False, True])
def test_secular_and_traces_vs_oracle(gpu_ctx, oracle_mod, gamma):
    s = configs.small_test_system(seed=9, gamma=gamma)
    print("[gpu] secular small gamma=%s" % gamma, flush=True)
    _setup(gpu_ctx, s)
    n = s.vale_dim
    coeffs = configs.pot_coefficients(s, seed=3)
    coeffs[1] = 0.0   # readPackedMatrixAccum: factor 0.0 adds unscaled
    ns = n - 2
    ev = gpu_ctx.secular(coeffs, ns)
    S = oracle_mod.assemble(s, "ol")
    K = oracle_mod.assemble(s, "ke")
    N = oracle_mod.assemble(s, "np")
    V = oracle_mod.assemble(s, "ep")
    occ = np.zeros((s.num_kpoints, ns))
    for k in range(s.num_kpoints):
        Hp = oracle_mod.build_hamiltonian(N[k], K[k], V[:, k], coeffs)
        H = oracle_mod.unpack(Hp, n)
        Sf = oracle_mod.unpack(S[k], n)
        w, vec, info = oracle_mod.hegv(H, Sf)
        assert info == 0
        assert np.max(np.abs(ev[k] - w[:ns])) < 1e-8, np.max(np.abs(ev[k] - w[:ns]))
        # eigenvectors: compare the projector onto the lowest 4 states
        X = gpu_ctx.eigenvectors(k, ns)[:, :4]
        Y = vec[:, :4]
        PX = X @ X.conj().T
        PY = Y @ Y.conj().T
        assert rel_err(PX, PY) < 1e-7
        # S-orthonormality of the GPU vectors (B-normalised like zhegv)
        G = X.conj().T @ Sf @ X
        assert np.allclose(G, np.eye(4), atol=1e-10)
    # occupations: fill 6 electrons with populateStandard
    occ = oracle_mod.populate(ev, s.kpoint_weights, 6)
    tr, prho = gpu_ctx.density_traces(occ)
    # oracle traces: rho(k) = sum f c c^H, then matrixElementMult summed over k
    t_ref = np.zeros(3)
    p_ref = np.zeros(s.pot_dim)
    for k in range(s.num_kpoints):
        X = gpu_ctx.eigenvectors(k, ns)
        rho = (X * occ[k]) @ X.conj().T
        iu = np.triu_indices(n)
        cols = np.concatenate([np.full(c + 1, c) for c in range(n)])
        rows = np.concatenate([np.arange(c + 1) for c in range(n)])
        rp = rho[rows, cols]
        if gamma:
            rp = rp.real
        del iu
        for q, M in enumerate((S[k], N[k], K[k])):
            t_ref[q] += oracle_mod.matrix_element_mult(M, rp, n)
        for j in range(s.pot_dim):
            p_ref[j] += oracle_mod.matrix_element_mult(V[j, k], rp, n)
    assert np.allclose(tr, t_ref, rtol=1e-10, atol=1e-12)
    assert np.allclose(prho, p_ref, rtol=1e-10, atol=1e-12)
    # Tr(rho S) = number of electrons
    assert tr[0] == pytest.approx(6.0, rel=1e-9)


def test_single_hegv_matches_lapack(gpu_ctx, oracle_mod):
    rng = np.random.default_rng(2)
    n = 40
    M = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = M + M.conj().T
    R = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = R @ R.conj().T + n * np.eye(n)
    w, V = gpu_ctx.hegv(H, S)
    w2, _, info = oracle_mod.hegv(H, S)
    assert np.max(np.abs(w - w2)) < 1e-10
    assert np.allclose(V.conj().T @ S @ V, np.eye(n), atol=1e-10)


@pytest.mark.parametrize("cx", [True, False])
@pytest.mark.parametrize("n", [7, 64, 150])
def test_single_hegv_reads_only_the_upper_triangle(gpu_ctx, oracle_mod, cx, n):
    """solveZHEGV / solveDSYGV's caller fills only the upper triangle of A
    (secularEqn.F90:140 zeroes valeVale, unpackMatrix fullFlag 0 writes rows
    <= column, matrixSubs.F90:426-433): A with a zero strictly lower
    triangle, and a nonzero imaginary part on the diagonal that zhegv 'U'
    ignores, must give LAPACK's eigenvalues of the Hermitian matrix the
    upper triangle defines."""
    rng = np.random.default_rng(40 + n)
    M = rng.normal(size=(n, n))
    R = rng.normal(size=(n, n))
    if cx:
        M = M + 1j * rng.normal(size=(n, n))
        R = R + 1j * rng.normal(size=(n, n))
    H = M + M.conj().T
    S = R @ R.conj().T + n * np.eye(n)
    Hu = np.triu(H)
    if cx:
        Hu[np.diag_indices(n)] += 1j * rng.normal(size=n)
    print("[gpu] hegv upper only n=%d cx=%s" % (n, cx), flush=True)
    w, V = gpu_ctx.hegv(Hu, S)
    w2, V2, info = oracle_mod.hegv(H, S)
    assert info == 0
    assert np.max(np.abs(w - w2)) < 1e-10 * max(1.0, np.max(np.abs(w2)))
    # the eigenvectors solve the full problem
    assert np.max(np.abs(H @ V - S @ V * w)) < 1e-9 * np.max(np.abs(H))
    assert np.allclose(V.conj().T @ S @ V, np.eye(n), atol=1e-10)


def test_not_positive_definite_overlap_is_reported(gpu_ctx):
    rng = np.random.default_rng(4)
    n = 16
    M = rng.normal(size=(n, n))
    H = M + M.T
    R = rng.normal(size=(n, 3))
    S = R @ R.T   # rank 3
    with pytest.raises(ImagoError) as e:
        gpu_ctx.hegv(H, S)
    assert e.value.name == "IMAGO_E_NOTPD"


@pytest.mark.parametrize("cx", [True, False])
@pytest.mark.parametrize("n", [2, 33, 97, 300])
def test_cholesky_factor_matches_lapack(gpu_ctx, cx, n):
    """The blocked Cholesky (32-column diagonal blocks, ragged last block)
    against LAPACK's zpotrf / dpotrf through scipy: B comes back from hegv
    with U (S = U^H U) in its upper triangle at 1e-12 relative, and its
    strictly lower triangle as it went in (zhegv's contract)."""
    import scipy.linalg as sl
    rng = np.random.default_rng(11 + n)
    R = rng.normal(size=(n, n))
    M = rng.normal(size=(n, n))
    if cx:
        R = R + 1j * rng.normal(size=(n, n))
        M = M + 1j * rng.normal(size=(n, n))
    S = R @ R.conj().T + n * np.eye(n)
    H = M + M.conj().T
    print("[gpu] cholesky n=%d cx=%s" % (n, cx), flush=True)
    _, _, got = gpu_ctx.hegv(H, S, return_factor=True)
    ref = sl.cholesky(S, lower=False)
    assert np.max(np.abs(np.triu(got) - ref)) < 1e-12 * np.max(np.abs(ref))
    lower = np.tril(np.ones((n, n), dtype=bool), -1)
    assert np.array_equal(got[lower], S[lower])


@pytest.mark.parametrize("cx", [True, False])
def test_not_positive_definite_reports_the_leading_minor(gpu_ctx, cx):
    """zpotrf's info through the blocked Cholesky: the first pivot that is
    not positive sits at order 71 (past the first diagonal blocks), so the error
    carries ZHEGV's code n + 71; a later negative pivot does not replace it."""
    n = 100
    rng = np.random.default_rng(6)
    S = np.eye(n) * 4.0
    E = rng.normal(size=(n, n)) * 0.01
    if cx:
        E = E + 1j * rng.normal(size=(n, n)) * 0.01
    S = S + E + E.conj().T
    S[70, 70] = -1.0
    S[90, 90] = -3.0
    H = np.diag(np.arange(n, dtype=float)).astype(S.dtype)
    with pytest.raises(ImagoError) as e:
        gpu_ctx.hegv(H, S)
    assert e.value.name == "IMAGO_E_NOTPD"
    assert "code=%d" % (n + 71) in str(e.value), str(e.value)


def test_secular_before_assembly_is_a_state_error(gpu_ctx):
    s = configs.small_test_system(seed=5)
    gpu_ctx.set_system(s)
    with pytest.raises(ImagoError) as e:
        gpu_ctx.secular(np.zeros(s.pot_dim), 4)
    assert e.value.name == "IMAGO_E_STATE"


def test_repeated_secular_reuses_overlap_factor(gpu_ctx, oracle_mod):
    """Successive secular calls on one assembled overlap (the SCF loop of
    mainSCF.F90) reuse the Cholesky factor of S(k); each must still match a
    fresh LAPACK solve of its own H(k), also after OL is reassembled."""
    s = configs.small_test_system(seed=11)
    print("[gpu] repeated secular", flush=True)
    _setup(gpu_ctx, s)
    n = s.vale_dim
    S = oracle_mod.assemble(s, "ol")
    K = oracle_mod.assemble(s, "ke")
    N = oracle_mod.assemble(s, "np")
    V = oracle_mod.assemble(s, "ep")
    for it, seed in enumerate((1, 2, 3)):
        if it == 2:
            gpu_ctx.assemble("ol")
        coeffs = configs.pot_coefficients(s, seed=seed)
        ev = gpu_ctx.secular(coeffs, n)
        for k in range(s.num_kpoints):
            H = oracle_mod.unpack(oracle_mod.build_hamiltonian(N[k], K[k], V[:, k], coeffs), n)
            w, _, info = oracle_mod.hegv(H, oracle_mod.unpack(S[k], n))
            assert info == 0
            assert np.max(np.abs(ev[k] - w)) < 1e-8, (it, k)
        # only the lowest states occupied: the density update stops there
        occ = np.zeros((s.num_kpoints, n))
        occ[:, :3] = 2.0 / s.num_kpoints
        tr, _ = gpu_ctx.density_traces(occ)
        ref = 0.0
        for k in range(s.num_kpoints):
            X = gpu_ctx.eigenvectors(k, n)
            rho = (X * occ[k]) @ X.conj().T
            Sf = oracle_mod.unpack(S[k], n)
            ref += np.real(np.sum(rho * Sf.T))
        assert tr[0] == pytest.approx(ref, rel=1e-10)
        # Tr(rho S) = sum of the occupations: 3 states x 2/nk at nk k-points
        assert tr[0] == pytest.approx(6.0, rel=1e-9)


def _secular_traces_oracle(oracle_mod, s, coeffs, occ_fn, ns):
    """Eigenvalues of every k and the traces (S, NP, KE, V_j) from the
    oracle's matrices for the GPU eigenvectors passed in via occ_fn."""
    n = s.vale_dim
    S = oracle_mod.assemble(s, "ol")
    K = oracle_mod.assemble(s, "ke")
    N = oracle_mod.assemble(s, "np")
    V = oracle_mod.assemble(s, "ep")
    evs = []
    for k in range(s.num_kpoints):
        H = oracle_mod.unpack(oracle_mod.build_hamiltonian(N[k], K[k], V[:, k], coeffs), n)
        w, _, info = oracle_mod.hegv(H, oracle_mod.unpack(S[k], n))
        assert info == 0
        evs.append(w[:ns])
    return np.array(evs), (S, N, K, V)


@pytest.mark.parametrize("name", ["small_complex", "small_gamma", "si8_cubic_d",
                                  "si2_primitive"])
def test_direct_ep_path_matches_oracle(gpu_ctx, oracle_mod, name):
    """IMAGO_EP_DIRECT: H(k) and Tr(rho V_j) from the stored primitive
    blocks instead of the potDim panels, against the oracle's panels; the
    Si2 primitive cell has 10 k-points (the H sum and the density fold take
    them four at a time)."""
    s = {"small_complex": lambda: configs.small_test_system(seed=9),
         "small_gamma": lambda: configs.small_test_system(seed=9, gamma=True),
         # cutoff 16 bohr: at 12 the screened overlap of this cell is not
         # positive definite
         "si8_cubic_d": lambda: configs.si64_supercell(reps=(1, 1, 1), mesh=(2, 2, 2),
                                                       cutoff_bohr=16.0),
         "si2_primitive": configs.si2_primitive}[name]()
    print("[gpu] direct EP", name, flush=True)
    gpu_ctx.set_system(s)
    for kind in ("ol", "ke", "np", "ep_direct"):
        gpu_ctx.assemble(kind)
    n = s.vale_dim
    coeffs = configs.pot_coefficients(s, seed=3)
    coeffs[1] = 0.0   # readPackedMatrixAccum: factor 0.0 adds unscaled
    ev = gpu_ctx.secular(coeffs, n)
    ref, (S, N, K, V) = _secular_traces_oracle(oracle_mod, s, coeffs, None, n)
    assert np.max(np.abs(ev - ref)) < 1e-8, np.max(np.abs(ev - ref))
    nel = 6 if name.startswith("small") else 14 * s.num_atoms
    occ = oracle_mod.populate(ev, s.kpoint_weights, nel)
    tr, prho = gpu_ctx.density_traces(occ)
    t_ref = np.zeros(3)
    p_ref = np.zeros(s.pot_dim)
    cols = np.concatenate([np.full(c + 1, c) for c in range(n)])
    rows = np.concatenate([np.arange(c + 1) for c in range(n)])
    for k in range(s.num_kpoints):
        X = gpu_ctx.eigenvectors(k, n)
        rp = ((X * occ[k]) @ X.conj().T)[rows, cols]
        if s.gamma:
            rp = rp.real
        for q, M in enumerate((S[k], N[k], K[k])):
            t_ref[q] += oracle_mod.matrix_element_mult(M, rp, n)
        for j in range(s.pot_dim):
            p_ref[j] += oracle_mod.matrix_element_mult(V[j, k], rp, n)
    assert np.allclose(tr, t_ref, rtol=1e-10, atol=1e-12)
    scale = np.max(np.abs(p_ref))
    assert np.max(np.abs(prho - p_ref)) <= 1e-10 * scale, np.max(np.abs(prho - p_ref)) / scale
    assert tr[0] == pytest.approx(nel, rel=1e-9)


@pytest.mark.parametrize("n", [2, 3, 33, 64, 97, 300])
def test_single_hegv_orders(gpu_ctx, oracle_mod, n):
    """The batched Hermitian eigensolver (tridiagonal reduction in 32-column
    panels, the panel-edge cases included) against LAPACK zhegv."""
    rng = np.random.default_rng(100 + n)
    M = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = M + M.conj().T
    R = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = R @ R.conj().T + n * np.eye(n)
    print("[gpu] hegv n=%d" % n, flush=True)
    w, V = gpu_ctx.hegv(H, S)
    w2, V2, info = oracle_mod.hegv(H, S)
    assert info == 0
    scale = max(1.0, np.max(np.abs(w2)))
    assert np.max(np.abs(w - w2)) < 1e-10 * scale, np.max(np.abs(w - w2))
    assert np.allclose(V.conj().T @ S @ V, np.eye(n), atol=1e-10)
    assert np.allclose(H @ V, S @ V * w, atol=1e-8 * scale)


@pytest.mark.parametrize("mode", ["coop", "rocsolver"])
@pytest.mark.parametrize("n", [2, 17, 97, 300, 1024])
def test_eigensolver_paths_orders(gpu_ctx, oracle_mod, monkeypatch, mode, n):
    """Both Hermitian eigensolvers against LAPACK zhegv up to the largest
    order of the cooperative one: the default (cooperative tridiagonal
    reduction, block-diagonal divide and conquer, compact-WY back-transform)
    and rocSOLVER's zheevd (IMAGO_HEEVD=rocsolver)."""
    monkeypatch.setenv("IMAGO_HEEVD", mode)
    rng = np.random.default_rng(300 + n)
    M = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = M + M.conj().T
    R = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = R @ R.conj().T + n * np.eye(n)
    print("[gpu] %s hegv n=%d" % (mode, n), flush=True)
    w, V = gpu_ctx.hegv(H, S)
    w2, _, info = oracle_mod.hegv(H, S)
    assert info == 0
    scale = max(1.0, np.max(np.abs(w2)))
    assert np.max(np.abs(w - w2)) < 1e-10 * scale, np.max(np.abs(w - w2))
    assert np.allclose(V.conj().T @ S @ V, np.eye(n), atol=1e-10)
    assert np.allclose(H @ V, S @ V * w, atol=1e-8 * scale)


def test_eigensolver_paths_batched_secular(gpu_ctx, monkeypatch):
    """A batch of k-points through the cooperative eigensolver gives
    rocSOLVER's eigenvalues and density traces."""
    s = configs.small_test_system(seed=12)
    n = s.vale_dim
    coeffs = configs.pot_coefficients(s, seed=5)
    occ = np.zeros((s.num_kpoints, n))
    occ[:, :4] = 2.0 / s.num_kpoints
    runs = []
    for mode in ("rocsolver", "coop"):
        monkeypatch.setenv("IMAGO_HEEVD", mode)
        print("[gpu] batched secular via", mode, flush=True)
        _setup(gpu_ctx, s)
        ev = gpu_ctx.secular(coeffs, n)
        tr, prho = gpu_ctx.density_traces(occ)
        runs.append((ev, tr, prho))
    (e1, t1, p1), (e2, t2, p2) = runs
    assert np.max(np.abs(e1 - e2)) < 1e-10 * max(1.0, np.max(np.abs(e1)))
    assert np.allclose(t1, t2, rtol=1e-10, atol=1e-12)
    assert np.max(np.abs(p1 - p2)) <= 1e-10 * max(1.0, np.max(np.abs(p1)))


@pytest.mark.parametrize("mode", ["coop", "rocsolver"])
@pytest.mark.parametrize("n", [2, 17, 300, 1024, 1500, 2048])
def test_real_eigensolver_paths_orders(gpu_ctx, oracle_mod, monkeypatch, mode, n):
    """Both real symmetric eigensolvers (the Gamma path) against LAPACK dsygv
    up to the largest order of the cooperative one."""
    monkeypatch.setenv("IMAGO_HEEVD", mode)
    rng = np.random.default_rng(500 + n)
    M = rng.normal(size=(n, n))
    H = M + M.T
    R = rng.normal(size=(n, n))
    S = R @ R.T + n * np.eye(n)
    print("[gpu] %s sygv n=%d" % (mode, n), flush=True)
    w, V = gpu_ctx.hegv(H, S)
    w2, _, info = oracle_mod.hegv(H, S)
    assert info == 0
    scale = max(1.0, np.max(np.abs(w2)))
    assert np.max(np.abs(w - w2)) < 1e-10 * scale, np.max(np.abs(w - w2))
    assert np.allclose(V.T @ S @ V, np.eye(n), atol=1e-10)
    assert np.allclose(H @ V, S @ V * w, atol=1e-8 * scale)


@pytest.mark.parametrize("gamma", [False, True])
def test_pending_columns_finish_on_demand(gpu_ctx, gamma):
    """imago_secular leaves the eigenvector columns to its consumers: the
    traces finish the occupied columns, imago_get_eigenvectors the rest.
    The vectors so assembled in two parts equal those of a secular call
    whose vectors are taken at once."""
    s = configs.small_test_system(seed=4, gamma=gamma)
    print("[gpu] pending columns gamma=%s" % gamma, flush=True)
    _setup(gpu_ctx, s)
    n = s.vale_dim
    coeffs = configs.pot_coefficients(s, seed=5)
    ev = gpu_ctx.secular(coeffs, n)
    occ = np.zeros((s.num_kpoints, n))
    occ[:, :3] = 2.0 / s.num_kpoints        # three occupied columns per k-point
    gpu_ctx.density_traces(occ)
    parts = [gpu_ctx.eigenvectors(k, n) for k in range(s.num_kpoints)]
    ev2 = gpu_ctx.secular(coeffs, n)
    whole = [gpu_ctx.eigenvectors(k, n) for k in range(s.num_kpoints)]
    assert np.array_equal(ev, ev2)
    for a, b in zip(parts, whole):
        assert rel_err(a, b) < 1e-12
