"""The oracle's restatement of the core orthogonalisation (oracle/ortho.py,
intgOrtho.F90 and orthoOL / ortho, integrals.F90:3608-3910) pinned by what
the mathematics fixes:

* the orthogonalised matrix is the operator between the projected valence
  functions |v'> = |v> - sum_c |c> <c|v>: X' = M^H X M with M = [I; -Scv],
  on random complex and real matrices and on assembled ones;
* with valence functions already orthogonal to the core ones nothing
  changes;
* the overlap's path (zherk with -2, then makeValeVale) equals the general
  operator path applied to the overlap itself;
* makeValeVale's flush sets components below 1e-10 to zero and keeps the
  rest;
* the extended matrices hold the valence-only ones at the positions
  extended_index gives."""
import numpy as np
import pytest

import ortho
from imago_amd import configs


def _pack_upper(full):
    n = full.shape[0]
    return np.concatenate([full[:j + 1, j] for j in range(n)])


def _random_ext(rng, n, c, cx):
    N = n + c
    A = rng.standard_normal((N, N)) + (1j * rng.standard_normal((N, N)) if cx else 0.0)
    S = A @ A.conj().T / N + np.eye(N)          # Hermitian positive definite
    B = rng.standard_normal((N, N)) + (1j * rng.standard_normal((N, N)) if cx else 0.0)
    X = B + B.conj().T                            # Hermitian
    perm = rng.permutation(N)                     # valence and core interleaved
    vale, core = np.sort(perm[:n]), np.sort(perm[n:])
    return S, X, vale, core


def _projected(X, S, vale, core):
    n = len(vale)
    M = np.zeros((X.shape[0], n), dtype=np.result_type(X, S))
    M[vale, np.arange(n)] = 1.0
    M[core, :] = -S[np.ix_(core, vale)]
    return M.conj().T @ X @ M


@pytest.mark.parametrize("cx", [True, False])
def test_operator_between_projected_functions(cx):
    rng = np.random.default_rng(1 if cx else 2)
    S, X, vale, core = _random_ext(rng, 9, 5, cx)
    want = _pack_upper(_projected(X, S, vale, core))
    got = ortho.ortho_op(X, S, vale, core)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    want_s = _pack_upper(_projected(S, S, vale, core))
    got_s = ortho.ortho_ol(S, vale, core)
    assert np.max(np.abs(got_s - want_s)) < 1e-12 * np.max(np.abs(want_s))
    # the overlap path is the operator path on the overlap
    assert np.max(np.abs(got_s - ortho.ortho_op(S, S, vale, core))) < 1e-13


@pytest.mark.parametrize("cx", [True, False])
def test_nothing_changes_without_core_overlap(cx):
    rng = np.random.default_rng(3)
    S, X, vale, core = _random_ext(rng, 7, 4, cx)
    S[np.ix_(core, vale)] = 0.0
    S[np.ix_(vale, core)] = 0.0
    X[np.abs(X) < 1e-9] = 1e-3   # keep the flush out of this check
    assert np.array_equal(ortho.ortho_op(X, S, vale, core),
                          _pack_upper(X[np.ix_(vale, vale)]))


def test_flush_below_1e10():
    vale, core = np.arange(3), np.arange(3, 4)
    X = np.zeros((4, 4), dtype=complex)
    X[0, 1] = 0.9e-10 + 2e-10j
    X[1, 0] = np.conj(X[0, 1])
    X[0, 2] = 1.1e-10 - 0.5e-10j
    X[2, 0] = np.conj(X[0, 2])
    X[1, 1] = -3.0
    S = np.eye(4, dtype=complex)
    p = ortho.ortho_op(X, S, vale, core)
    # packed upper: (0,0) (0,1) (1,1) (0,2) (1,2) (2,2)
    assert p[1] == 2e-10j
    assert p[3] == 1.1e-10
    assert p[2] == -3.0
    assert p[0] == 0.0


def test_extended_matrices_hold_the_valence_ones(oracle_mod):
    s = configs.small_test_system(seed=21, core=True)
    vale, core = ortho.extended_index(s)
    assert len(vale) == s.vale_dim and len(core) == s.core_dim
    N = s.vale_dim + s.core_dim
    for kind in ("ol", "np"):
        ext = oracle_mod.assemble(s.all_valence(), kind)
        base = oracle_mod.assemble(s, kind)
        for k in range(s.num_kpoints):
            E = ortho.unpack(ext[k], N)
            V = ortho.unpack(base[k], s.vale_dim)
            assert np.max(np.abs(E[np.ix_(vale, vale)] - V)) < 1e-13 * np.max(np.abs(V))


def test_orthogonalised_overlap_is_a_gram_matrix(oracle_mod):
    """On an assembled cell the orthogonalised overlap equals M^H S M and is
    positive definite."""
    s = configs.small_test_system(seed=22, core=True)
    vale, core = ortho.extended_index(s)
    N = s.vale_dim + s.core_dim
    ext = oracle_mod.assemble(s.all_valence(), "ol")
    got = ortho.orthogonalise(s, ext)
    for k in range(s.num_kpoints):
        S = ortho.unpack(ext[k], N)
        want = _projected(S, S, vale, core)
        full = ortho.unpack(got[k], s.vale_dim)
        assert np.max(np.abs(full - want)) < 1e-9
        assert np.linalg.eigvalsh(full).min() > 0.0
