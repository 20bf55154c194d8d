"""CPU restatement of the orthogonalisation of the valence functions
against the core functions (TEST INFRASTRUCTURE ONLY: the checker of the
library's core path; nothing in the product imports it).

* ``extended_index``  where each valence and core function sits in the
                      matrices of ``System.all_valence()`` (per atom its
                      valence, then its core states, as saveCurrentPair
                      reads currentPair, intgSaving.F90:346-455, 555-655)
* ``ortho_ol``        orthoOL for one k-point (integrals.F90:3608-3744):
                      valeCoreCoreValeOL (zherk 'U','C', alpha -2,
                      intgOrtho.F90:23-45), coreValeCoreCore (78-113) and
                      makeValeVale with storeFlag 1, fullFlag 0 (115-213:
                      the upper triangle, packed column by column, each
                      real and imaginary part below 1e-10 flushed to 0)
* ``ortho_op``        ortho for one k-point and operator (integrals.F90:
                      3747-3910): valeCoreCoreVale (zher2k 'U','C', alpha
                      -1, intgOrtho.F90:48-76), then coreValeCoreCore with
                      the operator's core-core block and makeValeVale
* ``orthogonalise``   both over the packed matrices the oracle assembles
                      for ``System.all_valence()``

The Gamma routines (intgOrtho.F90:215-393) are the same with real
matrices and transposes, and the flush of the one real component.
"""
from __future__ import annotations

import numpy as np

FLUSH = 1.0e-10   # smallThresh10, intgOrtho.F90:139


def extended_index(system):
    """(valence positions, core positions) in the all-valence ordering."""
    vale, core = [], []
    off = 0
    for t in system.atom_type_of:
        at = system.atom_types[t]
        nv, nc = at.num_states, at.num_core_states
        vale.extend(range(off, off + nv))
        core.extend(range(off + nv, off + nv + nc))
        off += nv + nc
    return np.array(vale, dtype=np.int64), np.array(core, dtype=np.int64)


def unpack(packed, n):
    """Full Hermitian n x n from the packed upper triangle (column by
    column; unpackMatrix with fullFlag 1, matrixSubs.F90:405-460)."""
    full = np.zeros((n, n), dtype=packed.dtype)
    q = 0
    for j in range(n):
        full[:j + 1, j] = packed[q:q + j + 1]
        q += j + 1
    full = full + np.triu(full, 1).conj().T
    return full


def _make_vale_vale(vv, vale_core_t, core_vale):
    """makeValeVale, storeFlag 1, fullFlag 0: vv(j,i) += sum_k
    valeCore(k,j) coreVale(k,i) for j <= i, packed with the flush."""
    n = vv.shape[0]
    upd = vale_core_t.T @ core_vale          # [j][i] = sum_k valeCore(k,j) coreVale(k,i)
    out = []
    for i in range(n):
        col = vv[:i + 1, i] + upd[:i + 1, i]
        out.append(col)
    p = np.concatenate(out)
    if np.iscomplexobj(p):
        re, im = p.real.copy(), p.imag.copy()
        re[np.abs(re) < FLUSH] = 0.0
        im[np.abs(im) < FLUSH] = 0.0
        return re + 1j * im
    p = p.copy()
    p[np.abs(p) < FLUSH] = 0.0
    return p


def _core_vale_core_core(core_vale, core_core):
    """coreValeCoreCore: valeCore(i,j) = sum_k conj(coreVale(k,j))
    coreCore(k,i), the (coreDim, valeDim) "pre-transposed" product."""
    return (core_vale.conj().T @ core_core).T


def ortho_ol(S_ext, vale, core):
    """orthoOL for one k-point: S_ext the full extended overlap."""
    Svv = S_ext[np.ix_(vale, vale)].copy()
    Scv = S_ext[np.ix_(core, vale)]
    Scc = S_ext[np.ix_(core, core)]
    Svv = Svv - 2.0 * (Scv.conj().T @ Scv)     # zherk, upper triangle used
    vc = _core_vale_core_core(Scv, Scc)
    return _make_vale_vale(Svv, vc, Scv)


def ortho_op(X_ext, S_ext, vale, core):
    """ortho for one k-point: X_ext the operator's full extended matrix,
    S_ext the overlap's (its core-valence block is coreValeOL)."""
    Xvv = X_ext[np.ix_(vale, vale)].copy()
    Xcv = X_ext[np.ix_(core, vale)]
    Xcc = X_ext[np.ix_(core, core)]
    Scv = S_ext[np.ix_(core, vale)]
    Xvv = Xvv - Scv.conj().T @ Xcv - Xcv.conj().T @ Scv   # zher2k, alpha -1
    vc = _core_vale_core_core(Scv, Xcc)
    return _make_vale_vale(Xvv, vc, Scv)


def orthogonalise(system, S_ext_packed, X_ext_packed=None):
    """Packed valence matrices per k-point: orthoOL of the extended overlap
    (X_ext_packed None) or ortho of an operator, from the packed matrices
    [nk][N(N+1)/2] assembled for system.all_valence()."""
    vale, core = extended_index(system)
    N = len(vale) + len(core)
    out = []
    for k in range(len(S_ext_packed)):
        S = unpack(np.asarray(S_ext_packed[k]), N)
        if X_ext_packed is None:
            out.append(ortho_ol(S, vale, core))
        else:
            out.append(ortho_op(unpack(np.asarray(X_ext_packed[k]), N), S, vale, core))
    return np.array(out)
