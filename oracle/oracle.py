"""Python handle on the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (the C restatement, oracle.c) and, when it has
been built here, oracle/_ref/libgauss_ref.so (the reference's own
primitive-integral module compiled from its Fortran sources).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.

The generalized eigensolve of the oracle is LAPACK's zhegv/dsygv through
scipy (itype=1, jobz='V', uplo='U'), exactly the call of solveZHEGV /
solveDSYGV (src/imago/interfaces.F90:73-200).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = ctypes.CDLL(path)
        for n in ("orc_overlap2c", "orc_kinetic2c"):
            getattr(L, n).argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp,
                                      ctypes.c_int, _dp]
        for n in ("orc_electron3c", "orc_nuclear3c"):
            getattr(L, n).argtypes = [ctypes.c_double] * 3 + [_dp, _dp, _dp,
                                                              ctypes.c_int, _dp]
        L.orc_boys.argtypes = [ctypes.c_double, _dp]
        L.orc_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, _dp]
        L.orc_assemble.restype = ctypes.c_int
        L.orc_last_tuple_count.restype = ctypes.c_long
        L.orc_find_lattice_vector.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_alpha_dist.argtypes = [ctypes.c_void_p, _dp, _dp,
                                     ctypes.POINTER(ctypes.c_int)]
        L.orc_matrix_element_mult.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_matrix_element_mult.restype = ctypes.c_double
        L.orc_use_external_primitives.argtypes = [ctypes.c_int]
        L.orc_error_fn.argtypes = [ctypes.c_double]
        L.orc_error_fn.restype = ctypes.c_double
        L.orc_elecstat.argtypes = [ctypes.c_void_p, ctypes.c_double] + [_dp] * 5
        L.orc_residual_q.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                     _dp]
        L.orc_residual_q_sites.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp,
                                           _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_xc_sample_vectors.argtypes = [ctypes.c_int, _dp]
        L.orc_xc_mesh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, _dp, _dp, _dp,
                                  ctypes.POINTER(ctypes.c_int)]
        L.orc_xc_mesh.restype = ctypes.c_int
        L.orc_xc_site.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, _dp, _dp,
                                  _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_xc_site.restype = ctypes.c_int
        L.orc_xc_overlap.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, _dp, _dp,
                                     _dp, _dp, ctypes.c_int, _dp]
        L.orc_xc_wigner.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_xc_potential.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, _dp,
                                       _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.orc_xc_overlap_sites.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, _dp,
                                           _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _dp]
        _lib = L
    return _lib


def ref_lib():
    """The reference's compiled primitive module, or None if not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libgauss_ref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        for n in ("ref_overlap2c", "ref_kinetic2c"):
            getattr(R, n).argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp,
                                      ctypes.c_int, _dp]
        for n in ("ref_electron3c", "ref_nuclear3c"):
            getattr(R, n).argtypes = [ctypes.c_double] * 3 + [_dp, _dp, _dp,
                                                              ctypes.c_int, _dp]
        R.ref_boys.argtypes = [ctypes.c_double, _dp]
        R.ref_error_fn.argtypes = [ctypes.c_double]
        R.ref_error_fn.restype = ctypes.c_double
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(_dp)


KINDS = {"ol": "overlap2c", "ke": "kinetic2c", "ep": "electron3c", "np": "nuclear3c"}


def primitive(kind, a, A, B, C, sw, source="oracle"):
    """One primitive block sh(16,16) (Fortran order) from the oracle
    restatement or from the reference module (source='ref')."""
    L = lib() if source == "oracle" else ref_lib()
    pre = "orc_" if source == "oracle" else "ref_"
    fn = getattr(L, pre + KINDS[kind])
    sh = np.zeros((16, 16), order="F")
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    if kind in ("ol", "ke"):
        fn(float(a[0]), float(a[1]), _p(A), _p(B), int(sw), _p(sh))
    else:
        C = np.ascontiguousarray(C, dtype=np.float64)
        fn(float(a[0]), float(a[1]), float(a[2]), _p(A), _p(B), _p(C), int(sw), _p(sh))
    return sh


def boys(T, source="oracle"):
    L = lib() if source == "oracle" else ref_lib()
    F = np.zeros(7)
    (L.orc_boys if source == "oracle" else L.ref_boys)(float(T), _p(F))
    return F


def use_reference_primitives(on: bool) -> bool:
    """Route the oracle's assembly loops through the reference's compiled
    primitives (oracle/_ref) when available.  Returns whether it is on."""
    L = lib()
    R = ref_lib()
    if on and R is None:
        return False
    if on:
        P2 = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_double, _dp, _dp,
                              ctypes.c_int, _dp)
        L.orc_set_external_primitives.argtypes = [ctypes.c_void_p] * 4
        addr = [ctypes.cast(getattr(R, n), ctypes.c_void_p).value
                for n in ("ref_overlap2c", "ref_kinetic2c", "ref_electron3c",
                          "ref_nuclear3c")]
        L.orc_set_external_primitives(*addr)
        del P2
    L.orc_use_external_primitives(1 if on else 0)
    return on


def assemble(system, kind, term_begin=0, term_end=None, raw=False):
    """orc_assemble: packed matrices [nterm][nk][npack] complex (or real for
    gamma) from the restatement of integrals.F90.  raw=True returns the
    reference's interleaved real(2, n(n+1)/2) layout as [nterm][nk][2*npack]."""
    from imago_amd._desc import make_desc
    kinds = {"ol": 1, "ke": 2, "np": 3, "ep": 4}
    k = kinds[kind]
    desc, keep = make_desc(system)
    n = system.vale_dim
    npack = n * (n + 1) // 2
    nk = system.num_kpoints
    if k != 4:
        term_begin, term_end = 0, 1
    elif term_end is None:
        term_end = system.pot_dim
    nterm = term_end - term_begin
    per = npack * (1 if system.gamma else 2) * nk
    out = np.zeros(nterm * per)
    rc = lib().orc_assemble(ctypes.byref(desc), k, term_begin, term_end, _p(out))
    if rc != 0:
        raise RuntimeError("oracle assembly failed with code %d" % rc)
    if raw:
        res = out.reshape(nterm, nk, -1)
    elif system.gamma:
        res = out.reshape(nterm, 1, npack)
    else:
        res = out.reshape(nterm, nk, npack, 2)
        res = res[..., 0] + 1j * res[..., 1]
    return res if kind == "ep" else res[0]


def error_fn(x, source="oracle"):
    """error_fn (mathSubs.f90:22-161) from the restatement or the
    reference's compiled module (source='ref')."""
    if source == "oracle":
        return lib().orc_error_fn(float(x))
    return ref_lib().ref_error_fn(float(x))


def elecstat(system, log_elec_thresh=None):
    """orc_elecstat (neutralAndNuclearQPot, elecStat.f90:59-664): the same
    dict as imago_amd.lib.Context.elecstat."""
    from imago_amd._desc import make_desc
    th = system.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
    desc, keep = make_desc(system)
    P = system.pot_dim
    mats = [np.zeros((P, P), order="F") for _ in range(3)]
    vecs = [np.zeros(P), np.zeros(P)]
    rc = lib().orc_elecstat(ctypes.byref(desc), float(th), *[_p(m) for m in mats],
                            *[_p(v) for v in vecs])
    if rc != 0:
        raise RuntimeError("oracle elecstat failed with code %d" % rc)
    return dict(potAlphaOverlap=mats[0], nonLocalNeutQPot=mats[1], localNeutQPot=mats[2],
                nonLocalNucQPot=vecs[0], localNucQPot=vecs[1])


def residual_q(system, log_elec_thresh=None, recip_cells=None, sites=None):
    """orc_residual_q (residualQ, elecStat.f90:665-1043): numPotTypes x
    potDim.  sites = (begin, end) restricts the outer site loop (the other
    columns stay 0)."""
    from imago_amd._desc import make_desc
    th = system.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
    dims, sizes = system.recip_cells(th) if recip_cells is None else recip_cells
    dims = np.ascontiguousarray(dims, dtype=np.float64)
    sizes = np.ascontiguousarray(sizes, dtype=np.float64)
    desc, keep = make_desc(system)
    out = np.zeros((len(system.pot_types), system.pot_dim), order="F")
    i0, i1 = (0, len(system.pot_site_pos)) if sites is None else sites
    rc = lib().orc_residual_q_sites(ctypes.byref(desc), float(th), len(sizes), _p(dims),
                                    _p(sizes), int(i0), int(i1), _p(out))
    if rc != 0:
        raise RuntimeError("oracle residual_q failed with code %d" % rc)
    return out


def xc_wigner(rho):
    """wignerXC (potentialUpdate.F90:2835-2867): (potential, energy density)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_xc_wigner(float(rho), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


class XcOracleMesh:
    """getECMeshParameters of the oracle for one system and mesh input."""

    def __init__(self, system, mesh=None, covalent_radius=None):
        from imago_amd._desc import make_desc, make_xc_params
        from imago_amd.system import XcMesh
        self.system = system
        self.mesh = mesh or XcMesh()
        self.params = make_xc_params(self.mesh)
        self.cov = np.ascontiguousarray(
            [p.covalent_radius for p in system.pot_types] if covalent_radius is None
            else covalent_radius, dtype=np.float64)
        self.desc, self._keep = make_desc(system)
        nv = self.mesh.num_sample_vectors
        ns = len(system.pot_site_pos)
        self.vec = np.zeros((nv, 3))
        lib().orc_xc_sample_vectors(nv, _p(self.vec))
        self.pls = np.zeros((ns, 3))
        self.maxrad = np.zeros((ns, nv))
        self.npts = np.zeros(ns, dtype=np.int32)
        self.maxpts = lib().orc_xc_mesh(ctypes.byref(self.desc), ctypes.byref(self.params),
                                        _p(self.cov), _p(self.vec), _p(self.pls),
                                        _p(self.maxrad),
                                        self.npts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))

    def site(self, site, gga=False):
        """(num_ray_points, radialWeight[maxpts], exchRhoOp[potDim, maxpts,
        numOp] in Fortran order)."""
        P = self.system.pot_dim
        nop = 10 if gga else 1
        rw = np.zeros(max(self.maxpts, 1))
        op = np.zeros((P, max(self.maxpts, 1), nop), order="F")
        n = lib().orc_xc_site(ctypes.byref(self.desc), ctypes.byref(self.params),
                              float(self.system.log_elec_thresh), _p(self.cov), _p(self.vec),
                              _p(self.pls), _p(self.maxrad), max(self.maxpts, 1), int(site),
                              1 if gga else 0, _p(rw), _p(op))
        return n, rw, op

    def potential(self, rho_total, rho_core=None):
        """exchCorrPot(:, 1:3) before the solve (Wigner, spin 1, LDA), as
        [3][potDim]."""
        P = self.system.pot_dim
        r1 = np.ascontiguousarray(rho_total, dtype=np.float64)
        r3 = np.zeros(P) if rho_core is None else np.ascontiguousarray(rho_core, dtype=np.float64)
        out = np.zeros((3, P))
        lib().orc_xc_potential(ctypes.byref(self.desc), ctypes.byref(self.params),
                               float(self.system.log_elec_thresh), _p(self.cov), _p(self.vec),
                               _p(self.pls), _p(self.maxrad), max(self.maxpts, 1), _p(r1),
                               _p(r3), _p(out))
        return out

    def overlap(self, sites=None):
        """exchCorrOverlap; sites = (begin, end) sums those sites only."""
        P = self.system.pot_dim
        out = np.zeros((P, P), order="F")
        i0, i1 = (0, len(self.system.pot_site_pos)) if sites is None else sites
        lib().orc_xc_overlap_sites(ctypes.byref(self.desc), ctypes.byref(self.params),
                                   float(self.system.log_elec_thresh), _p(self.cov),
                                   _p(self.vec), _p(self.pls), _p(self.maxrad),
                                   max(self.maxpts, 1), int(i0), int(i1), _p(out))
        return out


def last_tuple_count():
    return int(lib().orc_last_tuple_count())


def unpack(packed, n):
    """unpackMatrix fullFlag=1 (matrixSubs.F90:405-460): Hermitian full."""
    m = np.zeros((n, n), dtype=packed.dtype)
    iu = np.triu_indices(n)
    # packed order: column by column, rows 1..c
    cols = np.concatenate([np.full(c + 1, c) for c in range(n)])
    rows = np.concatenate([np.arange(c + 1) for c in range(n)])
    m[rows, cols] = packed
    m[cols, rows] = np.conj(packed)
    del iu
    return m


def hegv(H, S):
    """zhegv/dsygv itype=1, 'V', 'U' through scipy's LAPACK."""
    from scipy.linalg import lapack
    if np.iscomplexobj(H):
        w, v, info = lapack.zhegv(H, S, itype=1, jobz="V", uplo="U")
    else:
        w, v, info = lapack.dsygv(H, S, itype=1, jobz="V", uplo="U")
    return w, v, info


def build_hamiltonian(np_mat, ke_mat, v_terms, coeffs):
    """secularEqn.F90:148-171 with the factor==0 rule of
    readPackedMatrixAccum (matrixSubs.F90:369-373)."""
    h = np_mat + ke_mat
    for j, c in enumerate(coeffs):
        h = h + (v_terms[j] if c == 0.0 else v_terms[j] * c)
    return h


def hamiltonian(system, coeffs, chunk=32, threads=None):
    """H(k) = NP + KE + sum_j c_j V_j for every k-point, packed, with the
    potential terms assembled ``chunk`` at a time and accumulated in term
    order as secularEqnSCF's readPackedMatrixAccum loop does
    (secularEqn.F90:148-171, matrixSubs.F90:339-376: a factor of exactly
    0.0 adds the term unscaled); host memory stays at one chunk of panels.
    threads sets the assembly's OpenMP threads for the call."""
    if threads:
        lib().orc_set_threads(int(threads))
    try:
        h = assemble(system, "np", raw=True) + assemble(system, "ke", raw=True)   # [nk][..]
        P = system.pot_dim
        for j0 in range(0, P, chunk):
            V = assemble(system, "ep", j0, min(P, j0 + chunk), raw=True)
            for q in range(V.shape[0]):
                c = float(coeffs[j0 + q])
                if c == 0.0:
                    h += V[q]
                else:
                    h += V[q] * c
            del V
    finally:
        if threads:
            lib().orc_set_threads(1)
    if system.gamma:
        return h.reshape(1, -1)
    h = h.reshape(h.shape[0], -1, 2)
    return h[..., 0] + 1j * h[..., 1]


def build_hamiltonian_raw(np_mat, ke_mat, v_terms, coeffs):
    """orc_build_hamiltonian on interleaved packed arrays (C loop of
    secularEqn.F90:148-171)."""
    L = lib()
    L.orc_build_hamiltonian.argtypes = [ctypes.c_long, _dp, _dp, ctypes.c_int,
                                        ctypes.POINTER(_dp), _dp, _dp]
    n = len(np_mat)
    h = np.zeros(n)
    vs = [np.ascontiguousarray(v, dtype=np.float64) for v in v_terms]
    arr = (_dp * len(vs))(*[_p(v) for v in vs])
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    L.orc_build_hamiltonian(n, _p(np.ascontiguousarray(np_mat)), _p(np.ascontiguousarray(ke_mat)),
                            len(vs), arr, _p(c), _p(h))
    return h


def matrix_element_mult_raw(m1, m2, n, dim1=2):
    """matrixElementMult on interleaved packed arrays (no conversion)."""
    return lib().orc_matrix_element_mult(dim1, n, _p(m1), _p(m2))


def matrix_element_mult(m1, m2, n):
    """matrixElementMult (matrixSubs.F90:66-112) via the C restatement."""
    L = lib()
    if np.iscomplexobj(m1):
        a = np.ascontiguousarray(np.stack([m1.real, m1.imag], -1).ravel())
        b = np.ascontiguousarray(np.stack([m2.real, m2.imag], -1).ravel())
        return L.orc_matrix_element_mult(2, n, _p(a), _p(b))
    a = np.ascontiguousarray(m1, dtype=np.float64)
    b = np.ascontiguousarray(m2, dtype=np.float64)
    return L.orc_matrix_element_mult(1, n, _p(a), _p(b))


def populate(eigenvalues, kweights, num_electrons, spin=1):
    """populateStandard (populate.F90:190-485) for spin = 1: fill states in
    ascending energy with kPointWeight/spin electrons each; degenerate
    states at the Fermi level share the remainder.  eigenvalues [nk, ns].
    Returns occupations [nk, ns]."""
    nk, ns = eigenvalues.shape
    small = 1e-8
    flat = eigenvalues.reshape(-1)
    order = np.argsort(flat, kind="stable")
    pop = np.zeros(nk * ns)
    w = np.repeat(kweights / spin, ns)
    filled = 0.0
    occ_idx = len(order) - 1
    for i, idx in enumerate(order):
        pop[idx] = w[idx]
        filled += pop[idx]
        if filled > num_electrons:
            pop[idx] = w[idx] + num_electrons - filled
            filled = float(num_electrons)
        if abs(num_electrons - filled) < small:
            occ_idx = i
            break
    e_f = flat[order[occ_idx]]
    lo = hi = occ_idx
    degen = pop[order[occ_idx]]
    possible = w[order[occ_idx]]
    while hi + 1 < len(order) and abs(e_f - flat[order[hi + 1]]) <= small:
        hi += 1
        possible += w[order[hi]]
    while lo > 0 and abs(e_f - flat[order[lo - 1]]) <= small:
        lo -= 1
        possible += w[order[lo]]
        degen += w[order[lo]]
    for i in range(lo, hi + 1):
        pop[order[i]] = degen / possible * w[order[i]]
    return pop.reshape(nk, ns)
