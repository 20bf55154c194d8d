"""ctypes host interface to libimago_hip.so (include/imago_hip.h).

Mirrors the reference's call surface for the integral and solve stages:
``Context.assemble`` = gaussOverlapOL/KE/NP + elecPotGaussOverlap
(src/imago/integrals.F90), ``Context.secular`` = secularEqnSCF's H build and
zhegv/dsygv (secularEqn.F90:37-328, interfaces.F90:73-200),
``Context.density_traces`` = the traces of makeValenceRho
(valeCharge.F90:230-393).  Errors raise ``ImagoError`` carrying the
library's message, the counterpart of the reference's ``stop``.

There is no CPU fallback: if the HIP library is missing or no GPU is
visible, construction fails loudly.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._desc import make_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimago_hip.so")

OL, KE, NP, EP, EP_DIRECT = 1, 2, 3, 4, 5
KIND_CODES = {"ol": OL, "ke": KE, "np": NP, "ep": EP, "ep_direct": EP_DIRECT}

ERRORS = {
    -1: "IMAGO_E_ARG", -2: "IMAGO_E_HIP", -3: "IMAGO_E_LATTICE", -4: "IMAGO_E_NOTPD",
    -5: "IMAGO_E_SOLVE", -6: "IMAGO_E_STATE", -7: "IMAGO_E_SCREEN", -8: "IMAGO_E_NOMEM",
}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

EXPORTS = [
    "imago_ctx_create", "imago_ctx_destroy", "imago_last_error", "imago_set_system",
    "imago_primitive_batch", "imago_boys_batch", "imago_assemble", "imago_get_packed",
    "imago_secular", "imago_get_eigenvectors", "imago_density_traces", "imago_zhegv",
    "imago_dsygv", "imago_last_timing", "imago_ep_counts", "imago_elecstat", "imago_residual_q",
    "imago_xc_mesh", "imago_xc_rho_op", "imago_xc_overlap", "imago_xc_counts",
    "imago_xc_potential", "imago_pot_solve", "imago_populate", "imago_scf_setup",
    "imago_scf_update", "imago_site_descriptors", "imago_mlp_create", "imago_mlp_destroy",
    "imago_mlp_num_params", "imago_mlp_params", "imago_mlp_forward", "imago_mlp_gradient",
    "imago_mlp_adam_step", "imago_mlp_train", "imago_mlp_forward_noalloc", "imago_tune",
    "imago_tridiag_eig", "imago_dos", "imago_set_packed",
]


class ScfParams(ctypes.Structure):
    """imago_scf_params (the MAIN_INPUT_DATA block of imago.dat)."""
    _fields_ = [("num_electrons", ctypes.c_double), ("feedback_level", ctypes.c_int),
                ("relax_factor", ctypes.c_double), ("convg_test", ctypes.c_double),
                ("xc_code", ctypes.c_int), ("log_elec_thresh", ctypes.c_double)]


class ScfRecord(ctypes.Structure):
    """imago_scf_record (one fort.7 / fort.14 row)."""
    _fields_ = [("iteration", ctypes.c_int), ("converged", ctypes.c_int),
                ("electron_diff", ctypes.c_double), ("testable_delta", ctypes.c_double),
                ("kinetic_energy", ctypes.c_double), ("elec_stat_energy", ctypes.c_double),
                ("exch_corr_energy", ctypes.c_double), ("total_energy", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ImagoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code
        self.name = ERRORS.get(code, "?")


_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libimago_hip.so is not built: run __graft_entry__.build() "
                           "or python -m imago_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    L.imago_last_error.restype = ctypes.c_char_p
    L.imago_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.imago_ctx_destroy.argtypes = [ctypes.c_void_p]
    L.imago_set_system.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _ip, _ip]
    L.imago_primitive_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                        _ip, _dp]
    L.imago_boys_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.imago_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.imago_get_packed.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _dp]
    L.imago_set_packed.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _dp]
    L.imago_secular.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _ip]
    L.imago_get_eigenvectors.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.imago_density_traces.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.imago_zhegv.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _ip]
    L.imago_dsygv.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _ip]
    L.imago_last_timing.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.imago_ep_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long),
                                  ctypes.POINTER(ctypes.c_long), _ip,
                                  ctypes.POINTER(ctypes.c_long)]
    L.imago_elecstat.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]
    L.imago_residual_q.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                   _dp]
    L.imago_xc_mesh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, _dp, _ip]
    L.imago_xc_rho_op.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp]
    L.imago_xc_overlap.argtypes = [ctypes.c_void_p, _dp]
    L.imago_xc_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), _dp]
    L.imago_xc_potential.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.imago_pot_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _ip]
    L.imago_populate.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, _dp,
                                 _dp, _dp, _ip]
    L.imago_scf_setup.argtypes = [ctypes.c_void_p, ctypes.POINTER(ScfParams), ctypes.c_void_p,
                                  _dp, ctypes.c_int, _dp, _dp]
    L.imago_scf_update.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp,
                                   ctypes.POINTER(ScfRecord)]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.imago_site_descriptors.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                         _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _dp]
    L.imago_mlp_create.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_uint,
                                   ctypes.POINTER(ctypes.c_void_p)]
    L.imago_mlp_destroy.argtypes = [ctypes.c_void_p]
    L.imago_mlp_num_params.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
    L.imago_mlp_params.argtypes = [ctypes.c_void_p, _fp, _fp]
    L.imago_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp]
    L.imago_mlp_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp, _dp, _fp]
    L.imago_mlp_adam_step.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float]
    L.imago_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_uint, _dp]
    L.imago_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.imago_tridiag_eig.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                    _dp]
    L.imago_dos.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _ip, ctypes.c_int,
                            ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                            _dp, _dp, _dp]
    L.imago_mlp_forward_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise ImagoError(rc, load().imago_last_error().decode())


class Context:
    """One library context on one GPU (not thread-safe)."""

    def __init__(self, device: int = 0):
        L = load()
        h = ctypes.c_void_p()
        _check(L.imago_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.system = None
        self.vale_dim = 0
        self.pot_dim = 0

    def close(self):
        if self._h:
            load().imago_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up -------------------------------------------------------------
    def set_system(self, system):
        desc, keep = make_desc(system)
        n = ctypes.c_int()
        pd = ctypes.c_int()
        _check(load().imago_set_system(self._h, ctypes.byref(desc), ctypes.byref(n),
                                       ctypes.byref(pd)))
        del keep
        self.system = system
        self.vale_dim = n.value
        self.pot_dim = pd.value
        return self

    # -- primitives ---------------------------------------------------------
    def primitive_batch(self, kind, a, xyz, sw):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 9)
        sw = np.ascontiguousarray(sw, dtype=np.int32).reshape(-1)
        n = len(sw)
        out = np.zeros((n, 256))
        _check(load().imago_primitive_batch(self._h, KIND_CODES.get(kind, kind), n, _p(a),
                                            _p(xyz), sw.ctypes.data_as(_ip), _p(out)))
        return out.reshape(n, 16, 16).transpose(0, 2, 1)   # sh[t][i][j] = sh(i,j)

    def boys_batch(self, T):
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
        F = np.zeros((len(T), 7))
        _check(load().imago_boys_batch(self._h, len(T), _p(T), _p(F)))
        return F

    # -- assembly -----------------------------------------------------------
    def assemble(self, kind):
        _check(load().imago_assemble(self._h, KIND_CODES.get(kind, kind)))
        return self

    def get_packed(self, kind, term=0, kpoint=0):
        n = self.vale_dim
        npack = n * (n + 1) // 2
        if self.system.gamma:
            out = np.zeros(npack)
            _check(load().imago_get_packed(self._h, KIND_CODES.get(kind, kind), term, kpoint,
                                           _p(out)))
            return out
        out = np.zeros(2 * npack)
        _check(load().imago_get_packed(self._h, KIND_CODES.get(kind, kind), term, kpoint,
                                       _p(out)))
        return out[0::2] + 1j * out[1::2]

    def get_packed_raw(self, kind, term=0, kpoint=0):
        """The packed matrix as the C layout holds it: real and imaginary
        parts interleaved (complex build) or real (gamma)."""
        n = self.vale_dim
        out = np.zeros((1 if self.system.gamma else 2) * (n * (n + 1) // 2))
        _check(load().imago_get_packed(self._h, KIND_CODES.get(kind, kind), term, kpoint,
                                       _p(out)))
        return out

    def set_packed(self, kind, packed, term=0, kpoint=0):
        """Load one packed matrix (complex values, or the interleaved /
        real layout of get_packed_raw) in place of assembling it."""
        n = self.vale_dim
        npack = n * (n + 1) // 2
        a = np.asarray(packed)
        if np.iscomplexobj(a):
            if self.system.gamma:
                raise ValueError("a gamma build takes real packed matrices")
            buf = np.empty(2 * npack)
            buf[0::2], buf[1::2] = a.real, a.imag
        else:
            buf = np.ascontiguousarray(a, dtype=np.float64)
        if buf.size != (1 if self.system.gamma else 2) * npack:
            raise ValueError("packed matrix of %d doubles for valeDim %d" % (buf.size, n))
        _check(load().imago_set_packed(self._h, KIND_CODES.get(kind, kind), term, kpoint,
                                       _p(buf)))
        return self

    # -- secular equation ---------------------------------------------------
    def secular(self, pot_coeffs, num_states):
        nk = self.system.num_kpoints
        c = np.ascontiguousarray(pot_coeffs, dtype=np.float64)
        ev = np.zeros(nk * num_states)
        info = np.zeros(nk, dtype=np.int32)
        rc = load().imago_secular(self._h, _p(c), int(num_states), _p(ev),
                                  info.ctypes.data_as(_ip))
        self.last_info = info
        _check(rc)
        return ev.reshape(nk, num_states)

    def eigenvectors(self, kpoint, num_states):
        n = self.vale_dim
        if self.system.gamma:
            out = np.zeros(n * num_states)
            _check(load().imago_get_eigenvectors(self._h, kpoint, _p(out)))
            return out.reshape(num_states, n).T
        out = np.zeros(2 * n * num_states)
        _check(load().imago_get_eigenvectors(self._h, kpoint, _p(out)))
        z = out[0::2] + 1j * out[1::2]
        return z.reshape(num_states, n).T

    def density_traces(self, occupations):
        occ = np.ascontiguousarray(occupations, dtype=np.float64)
        tr = np.zeros(3)
        pr = np.zeros(max(self.pot_dim, 1))
        _check(load().imago_density_traces(self._h, _p(occ), _p(tr), _p(pr)))
        return tr, pr[:self.pot_dim]

    def timing(self, stage):
        v = ctypes.c_double()
        _check(load().imago_last_timing(self._h, stage, ctypes.byref(v)))
        return v.value

    def ep_counts(self):
        """(tuples, items, doubles per primitive block, doubles of the direct
        H-sum run partials) of the EP screening."""
        t, i, b, w = ctypes.c_long(), ctypes.c_long(), ctypes.c_int(), ctypes.c_long()
        _check(load().imago_ep_counts(self._h, ctypes.byref(t), ctypes.byref(i), ctypes.byref(b),
                                      ctypes.byref(w)))
        return t.value, i.value, b.value, w.value

    # -- electrostatics (elecStat.f90) --------------------------------------
    def elecstat(self, log_elec_thresh=None):
        """neutralAndNuclearQPot of the current system: dict with
        potAlphaOverlap, nonLocalNeutQPot, localNeutQPot (potDim x potDim,
        [r, c] = (term of site j, term of site i) as the reference's arrays),
        nonLocalNucQPot and localNucQPot (potDim)."""
        th = self.system.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
        P = self.pot_dim
        mats = [np.zeros((P, P), order="F") for _ in range(3)]
        vecs = [np.zeros(P), np.zeros(P)]
        _check(load().imago_elecstat(self._h, float(th), *[_p(m) for m in mats],
                                     *[_p(v) for v in vecs]))
        return dict(potAlphaOverlap=mats[0], nonLocalNeutQPot=mats[1], localNeutQPot=mats[2],
                    nonLocalNucQPot=vecs[0], localNucQPot=vecs[1])

    def residual_q(self, log_elec_thresh=None, recip_cells=None):
        """residualQ of the current system: numPotTypes x potDim
        nonLocalResidualQ.  recip_cells = (dims [n][3], sizes [n]); the
        system's reciprocal lattice by default."""
        th = self.system.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
        dims, sizes = (self.system.recip_cells(th) if recip_cells is None else recip_cells)
        dims = np.ascontiguousarray(dims, dtype=np.float64)
        sizes = np.ascontiguousarray(sizes, dtype=np.float64)
        out = np.zeros((len(self.system.pot_types), self.pot_dim), order="F")
        _check(load().imago_residual_q(self._h, float(th), len(sizes), _p(dims), _p(sizes),
                                       _p(out)))
        return out

    # -- exchange-correlation mesh (exchCorr.f90) ---------------------------
    def xc_mesh(self, mesh=None, log_elec_thresh=None, covalent_radius=None):
        """getECMeshParameters for the current system; returns
        maxNumRayPoints.  mesh: an imago_amd.system.XcMesh (makeinput's
        defaults when None); covalent radii from the potential types."""
        from imago_amd._desc import make_xc_params
        from imago_amd.system import XcMesh
        self._xc_mesh = mesh or XcMesh()
        params = make_xc_params(self._xc_mesh)
        th = self.system.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
        cov = np.ascontiguousarray(
            [p.covalent_radius for p in self.system.pot_types] if covalent_radius is None
            else covalent_radius, dtype=np.float64)
        mx = ctypes.c_int()
        _check(load().imago_xc_mesh(self._h, ctypes.byref(params), float(th), _p(cov),
                                    ctypes.byref(mx)))
        self.xc_maxpts = mx.value
        return mx.value

    def xc_rho_op(self, site, gga=False):
        """(num_ray_points, radialWeight[maxpts], exchRhoOp[potDim, maxpts,
        numOp] in Fortran order) of one site."""
        mp = max(self.xc_maxpts, 1)
        nop = 10 if gga else 1
        rw = np.zeros(mp)
        op = np.zeros((self.pot_dim, mp, nop), order="F")
        n = ctypes.c_int()
        _check(load().imago_xc_rho_op(self._h, int(site), 1 if gga else 0, ctypes.byref(n),
                                      _p(rw), _p(op)))
        return n.value, rw, op

    def xc_overlap(self):
        """exchCorrOverlap, potDim x potDim (upper triangle)."""
        out = np.zeros((self.pot_dim, self.pot_dim), order="F")
        _check(load().imago_xc_overlap(self._h, _p(out)))
        return out

    def xc_potential(self, rho_total, rho_core=None):
        """exchCorrPot(:, 1:3) before the solve (Wigner, spin 1, LDA), as
        [3][potDim]."""
        r1 = np.ascontiguousarray(rho_total, dtype=np.float64)
        r3 = None if rho_core is None else np.ascontiguousarray(rho_core, dtype=np.float64)
        out = np.zeros((3, self.pot_dim))
        _check(load().imago_xc_potential(self._h, _p(r1), None if r3 is None else _p(r3), _p(out)))
        return out

    def pot_solve(self, matrix, b):
        """solveDPOSVX with potAlphaOverlap (matrix='pot') or exchCorrOverlap
        ('xc') on the device; b [potDim] or [potDim, nrhs]."""
        which = {"pot": 1, "xc": 2}[matrix]
        x = np.asfortranarray(np.array(b, dtype=np.float64, copy=True))
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        info = ctypes.c_int()
        _check(load().imago_pot_solve(self._h, which, nrhs, _p(x), ctypes.byref(info)))
        return x

    # -- SCF (makeSCFPot) -----------------------------------------------------
    def scf_setup(self, num_electrons, mesh=None, feedback_level=2, relax_factor=0.2,
                  convg_test=1.0e-4, xc_code=100, log_elec_thresh=None, covalent_radius=None,
                  recip_cells=None):
        """imago_scf_setup for the current system: the electrostatic and
        exchange-correlation operators of makeSCFPot on the device."""
        from imago_amd._desc import make_xc_params
        from imago_amd.system import XcMesh
        th = self.system.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
        params = ScfParams(float(num_electrons), int(feedback_level), float(relax_factor),
                           float(convg_test), int(xc_code), float(th))
        xcp = make_xc_params(mesh or XcMesh())
        cov = np.ascontiguousarray(
            [p.covalent_radius for p in self.system.pot_types] if covalent_radius is None
            else covalent_radius, dtype=np.float64)
        dims, sizes = (self.system.recip_cells(th) if recip_cells is None else recip_cells)
        dims = np.ascontiguousarray(dims, dtype=np.float64)
        sizes = np.ascontiguousarray(sizes, dtype=np.float64)
        _check(load().imago_scf_setup(self._h, ctypes.byref(params), ctypes.byref(xcp), _p(cov),
                                      len(sizes), _p(dims), _p(sizes)))

    def scf_update(self, traces3, pot_rho, pot_coeffs):
        """imago_scf_update (makeSCFPot): returns (next coefficients,
        record dict, fitted charge [2, potDim])."""
        tr = np.ascontiguousarray(traces3, dtype=np.float64)
        pr = np.ascontiguousarray(pot_rho, dtype=np.float64)
        c = np.array(pot_coeffs, dtype=np.float64, copy=True)
        rho = np.zeros((2, self.pot_dim))
        rec = ScfRecord()
        _check(load().imago_scf_update(self._h, _p(tr), _p(pr), _p(c), _p(rho),
                                       ctypes.byref(rec)))
        return c, rec.as_dict(), rho

    TUNE = {"potrf_block": 1, "eigensolver": 2, "cholesky": 3, "tridiag": 4, "graph": 5}

    def tune(self, **kw):
        """imago_tune: potrf_block=1..32, eigensolver=0|1, cholesky=0|1,
        tridiag=0|1 (the divide and conquer: this library's or rocSOLVER's),
        graph=0|1 (the divide and conquer replayed from a captured graph)."""
        for k, v in kw.items():
            _check(load().imago_tune(self._h, self.TUNE[k], int(v)))

    def tridiag_eig(self, d, e):
        """imago_tridiag_eig: eigenvalues (ascending) and eigenvectors
        (columns) of symmetric tridiagonals; d (nb, n) or (n,), e (nb, n-1)
        or (n-1,).  Returns (w, Z) of shapes (nb, n), (nb, n, n)."""
        d = np.ascontiguousarray(np.atleast_2d(d), dtype=np.float64)
        nb, n = d.shape
        e = np.ascontiguousarray(np.asarray(e, dtype=np.float64).reshape(nb, max(n - 1, 0)))
        w = np.empty((nb, n))
        Z = np.empty((nb, n, n))
        ep = e if e.size else np.zeros(1)
        _check(load().imago_tridiag_eig(self._h, n, nb, _p(d), _p(ep), _p(w), _p(Z)))
        return w, np.ascontiguousarray(np.swapaxes(Z, 1, 2))

    def dos(self, energies, spin, channel, num_channels, emin, delta, num_points, sigma):
        """imago_dos on the eigenpairs of the last secular call: returns
        (pdos (num_channels, num_points), electron number per basis
        function, localisation index per state)."""
        e = np.ascontiguousarray(energies, dtype=np.float64)
        ch = np.ascontiguousarray(channel, dtype=np.int32)
        pd = np.zeros((num_points, num_channels))
        en = np.zeros(self.vale_dim)
        loc = np.zeros(e.shape[-1])
        _check(load().imago_dos(self._h, _p(e), int(spin), ch.ctypes.data_as(_ip),
                                int(num_channels), float(emin), float(delta), int(num_points),
                                float(sigma), _p(pd), _p(en), _p(loc)))
        return pd.T.copy(), en, loc

    def xc_counts(self):
        """(all ray points, flops of the last overlap's GEMMs)."""
        n, f = ctypes.c_long(), ctypes.c_double()
        _check(load().imago_xc_counts(self._h, ctypes.byref(n), ctypes.byref(f)))
        return n.value, f.value

    # -- single problems ----------------------------------------------------
    def hegv(self, H, S, return_factor=False):
        """zhegv/dsygv itype=1 'V' 'U' on the GPU: returns (w, vectors), and
        with return_factor B as zhegv leaves it (U of S = U^H U in the upper
        triangle, the strictly lower triangle of S unchanged)."""
        n = H.shape[0]
        info = ctypes.c_int()
        w = np.zeros(n)
        if np.iscomplexobj(H):
            A = np.asfortranarray(H, dtype=np.complex128).copy(order="F")
            B = np.asfortranarray(S, dtype=np.complex128).copy(order="F")
            _check(load().imago_zhegv(self._h, n, A.ctypes.data_as(_dp), B.ctypes.data_as(_dp),
                                      _p(w), ctypes.byref(info)))
        else:
            A = np.asfortranarray(H, dtype=np.float64).copy(order="F")
            B = np.asfortranarray(S, dtype=np.float64).copy(order="F")
            _check(load().imago_dsygv(self._h, n, A.ctypes.data_as(_dp), B.ctypes.data_as(_dp),
                                      _p(w), ctypes.byref(info)))
        return (w, A, B) if return_factor else (w, A)


def populate(eigenvalues, kpoint_weights, num_electrons):
    """imago_populate (populateStandard): occupations [nk, ns] and
    (occupied energy, band gap, gap kind)."""
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    nk, ns = ev.shape
    kw = np.ascontiguousarray(kpoint_weights, dtype=np.float64)
    pop = np.zeros((nk, ns))
    occ, gap = ctypes.c_double(), ctypes.c_double()
    kind = ctypes.c_int()
    _check(load().imago_populate(nk, ns, _p(ev), _p(kw), float(num_electrons), _p(pop),
                                 ctypes.byref(occ), ctypes.byref(gap), ctypes.byref(kind)))
    return pop, (occ.value, gap.value, kind.value)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """imago_mlp: the per-site MLP on a context's device (FP32)."""

    def __init__(self, ctx, widths, seed=0):
        self.widths = [int(w) for w in widths]
        h = ctypes.c_void_p()
        w = np.ascontiguousarray(self.widths, dtype=np.int32)
        _check(load().imago_mlp_create(ctx._h, len(w) - 1, w.ctypes.data_as(_ip), int(seed),
                                       ctypes.byref(h)))
        self._h = h
        n = ctypes.c_long()
        _check(load().imago_mlp_num_params(self._h, ctypes.byref(n)))
        self.num_params = n.value

    def close(self):
        if self._h:
            load().imago_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        _check(load().imago_mlp_params(self._h, _fptr(p), None))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(load().imago_mlp_params(self._h, None, _fptr(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.widths[-1]), dtype=np.float32)
        _check(load().imago_mlp_forward(self._h, x.shape[0], _fptr(x), _fptr(y)))
        return y

    def gradient(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_double()
        _check(load().imago_mlp_gradient(self._h, x.shape[0], _fptr(x), _fptr(t),
                                         ctypes.byref(loss), _fptr(g)))
        return loss.value, g

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(load().imago_mlp_adam_step(self._h, lr, beta1, beta2, eps))

    def train(self, x, target, epochs, minibatch=256, lr=1e-3, beta1=0.9, beta2=0.999,
              eps=1e-8, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        losses = np.zeros(epochs)
        _check(load().imago_mlp_train(self._h, x.shape[0], _fptr(x), _fptr(t), int(epochs),
                                      int(minibatch), lr, beta1, beta2, eps, int(seed),
                                      _p(losses)))
        return losses


def site_descriptors(ctx, pos, weight, cells, cutoff, r_min, eta):
    """imago_site_descriptors: [num_sites][96]."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    w = np.ascontiguousarray(weight, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    out = np.zeros((len(pos), 96))
    _check(load().imago_site_descriptors(ctx._h, len(pos), _p(pos), _p(w), len(cells),
                                         _p(cells), float(cutoff), float(r_min), float(eta),
                                         _p(out)))
    return out
