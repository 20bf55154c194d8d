"""Symmetric tridiagonal test matrices for the divide and conquer: random,
Wilkinson's W21+ and glued copies of it (clusters to 1e-10), constant
diagonals with tiny off-diagonals (a 50-fold cluster), zero off-diagonals
(the problem splits), graded and reverse-graded entries, the 1-2-1
Toeplitz matrix (closed-form spectrum), alternating signs, repeated
diagonal entries, a spectrum with 5-fold clusters from a Householder
tridiagonalisation, and scalings by 1e8 and 1e-9."""
import numpy as np


def cases(max_n=300, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for n in (1, 2, 3, 4, 5, 7, 16, 17, 33, 100, 257):
        if n <= max_n:
            out.append(("rand%d" % n, rng.normal(size=n), rng.normal(size=n - 1)))
    w = np.abs(np.arange(21) - 10).astype(float)
    out.append(("wilkinson21", w, np.ones(20)))
    e = np.ones(5 * 21 - 1)
    e[20::21] = 1e-10
    out.append(("glued_wilkinson", np.tile(w, 5), e))
    out.append(("constant", np.full(50, 2.0), np.full(49, 1e-14)))
    out.append(("zero_offdiag", rng.normal(size=40), np.zeros(39)))
    e = rng.normal(size=59)
    e[::7] = 0.0
    out.append(("split", rng.normal(size=60), e))
    out.append(("graded", 10.0 ** (-np.arange(30) / 3.0), 10.0 ** (-np.arange(29) / 3.0 - 0.2)))
    out.append(("reverse_graded", 10.0 ** (np.arange(40) / 4.0),
                10.0 ** (np.arange(39) / 4.0 - 0.5)))
    out.append(("toeplitz", np.full(64, 2.0), np.full(63, -1.0)))
    out.append(("alternating", np.zeros(128), np.where(np.arange(127) % 2, 1.0, -1.0)))
    out.append(("repeated_diag", np.repeat(rng.normal(size=20), 8), rng.normal(size=159) * 1e-6))
    out.append(("scaled_1e8", rng.normal(size=120) * 1e8, rng.normal(size=119) * 1e3))
    out.append(("scaled_1e-9", rng.normal(size=120) * 1e-9, rng.normal(size=119) * 1e-12))
    n = min(200, max_n)
    lam = np.sort(np.concatenate([np.repeat(rng.normal(size=n // 10), 5),
                                  rng.normal(size=n - 5 * (n // 10)) * 100]))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = (Q * lam) @ Q.T
    d, e = _householder_tridiagonal(A)
    out.append(("clustered", d, e))
    return out


def _householder_tridiagonal(A):
    """Symmetric A reduced to tridiagonal form by Householder reflections."""
    A = np.array(A, dtype=float)
    n = A.shape[0]
    for k in range(n - 2):
        x = A[k + 1:, k].copy()
        alpha = -np.copysign(np.linalg.norm(x), x[0])
        v = x.copy()
        v[0] -= alpha
        nv = np.linalg.norm(v)
        if nv == 0.0:
            continue
        v /= nv
        A[k + 1:, :] -= 2.0 * np.outer(v, v @ A[k + 1:, :])
        A[:, k + 1:] -= 2.0 * np.outer(A[:, k + 1:] @ v, v)
    return np.diag(A).copy(), np.diag(A, -1).copy()


def check(d, e, w, Z):
    """(eigenvalue error vs LAPACK, orthogonality, residual), the first and
    last relative to max |T|."""
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    nrm = max(np.abs(T).max(), 1e-300)
    wr = np.linalg.eigvalsh(T)
    return (np.abs(w - wr).max() / nrm, np.abs(Z.T @ Z - np.eye(n)).max(),
            np.abs(T @ Z - Z * w).max() / nrm)
