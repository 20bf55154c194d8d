"""Checkpoint and resume of the integral stage in the reference's SCF file.

The reference keeps every integral matrix it computes in an HDF5 file
(initHDF5_SCF, src/imago/hdf5SCF.F90:60-230: "gs_scf-fb.hdf5" for the
ground state in the full basis, a group per k-point set such as
"00004_00004_00004", under it atomIntgGroup; hdf5SCFIntg.F90:105-525) and
marks each integral group with a "status" attribute once all of its
k-points are written (integrals.F90:3730, 3876, 4029).  A later run on the
same file and k-point set reads the matrices of every group whose status is
1 instead of computing them (integrals.F90:262-294 and the matching checks
of the KE, NP and potential-term routines).

``IntegralFile`` is that file through libimago_h5.so (csrc/h5store.cpp);
``assemble_or_load`` is the resume: each kind read from the file when its
status is 1, else assembled on the device, written and marked.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import lib

HERE = os.path.dirname(os.path.abspath(__file__))
H5LIB = os.path.join(HERE, "libimago_h5.so")

# the file's integral groups; OL .. EP share the GPU library's kind codes
GROUP_CODES = {"ol": 1, "ke": 2, "np": 3, "ep": 4, "mv": 5}

_L = None


def load_h5():
    """libimago_h5.so, built by imago_amd.build (needs the HDF5 C library)."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(H5LIB):
        raise RuntimeError("libimago_h5.so is not built (python -m imago_amd.build; "
                           "it needs hdf5.h and libhdf5)")
    L = ctypes.CDLL(H5LIB)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.imago_h5_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ip, ip, ip, ip,
                                ctypes.POINTER(vp), ctypes.POINTER(ip)]
    L.imago_h5_close.argtypes = [vp]
    L.imago_h5_write_packed.argtypes = [vp, ip, ip, ip, dp]
    L.imago_h5_read_packed.argtypes = [vp, ip, ip, ip, dp]
    L.imago_h5_status.argtypes = [vp, ip, ip, ctypes.POINTER(ip)]
    L.imago_h5_set_status.argtypes = [vp, ip, ip, ip]
    L.imago_h5_last_error.restype = ctypes.c_char_p
    _L = L
    return L


def _h5check(rc):
    if rc != 0:
        raise RuntimeError("imago_h5: " + load_h5().imago_h5_last_error().decode())


def file_name(basis="fb", edge="gs"):
    """The reference's file name (hdf5SCF.F90:83-105): edge "gs" for the
    ground state (or e.g. "1s" for a core hole), basis "mb", "fb" or "eb"."""
    return "%2s_scf-%s.hdf5" % (edge, basis)


def kpoint_group(mesh):
    """The k-point set's group name for an axial mesh (hdf5SCF.F90:132-133)."""
    a, b, c = (int(x) for x in mesh)
    return "%05d_%05d_%05d" % (a, b, c)


class IntegralFile:
    """The integral part of the reference's SCF file for one k-point set."""

    def __init__(self, path, group, num_kpoints, vale_dim, pot_dim, gamma=False):
        L = load_h5()
        self._h = ctypes.c_void_p()
        created = ctypes.c_int(0)
        _h5check(L.imago_h5_open(os.fsencode(path), group.encode(), num_kpoints, vale_dim,
                                 pot_dim, 1 if gamma else 0, ctypes.byref(self._h),
                                 ctypes.byref(created)))
        self.created = bool(created.value)
        self.num_kpoints, self.vale_dim, self.pot_dim = num_kpoints, vale_dim, pot_dim
        self.gamma = bool(gamma)
        self.packed_len = (1 if gamma else 2) * (vale_dim * (vale_dim + 1) // 2)

    @classmethod
    def for_system(cls, path, system, mesh):
        """The file of a system whose k-points come from the axial mesh."""
        return cls(path, kpoint_group(mesh), system.num_kpoints, system.vale_dim,
                   system.pot_dim, system.gamma)

    def close(self):
        if self._h:
            load_h5().imago_h5_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def status(self, kind, term=0):
        v = ctypes.c_int(0)
        _h5check(load_h5().imago_h5_status(self._h, GROUP_CODES.get(kind, kind), term,
                                           ctypes.byref(v)))
        return v.value

    def set_status(self, kind, term=0, value=1):
        _h5check(load_h5().imago_h5_set_status(self._h, GROUP_CODES.get(kind, kind), term,
                                               value))

    def write(self, kind, packed, term=0, kpoint=0):
        buf = np.ascontiguousarray(packed, dtype=np.float64)
        if buf.size != self.packed_len:
            raise ValueError("packed matrix of %d doubles, the file holds %d"
                             % (buf.size, self.packed_len))
        _h5check(load_h5().imago_h5_write_packed(
            self._h, GROUP_CODES.get(kind, kind), term, kpoint,
            buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def read(self, kind, term=0, kpoint=0):
        out = np.zeros(self.packed_len)
        _h5check(load_h5().imago_h5_read_packed(
            self._h, GROUP_CODES.get(kind, kind), term, kpoint,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out


def _terms(f: IntegralFile, kind):
    return range(f.pot_dim) if kind == "ep" else range(1)


def save_integrals(ctx: lib.Context, f: IntegralFile, kinds=("ol", "ke", "np", "ep")):
    """Write the context's assembled matrices of each kind to the file and
    mark each group complete, as the reference's ortho routines do after
    their last k-point (integrals.F90:3868-3886)."""
    for kind in kinds:
        for term in _terms(f, kind):
            for k in range(f.num_kpoints):
                f.write(kind, ctx.get_packed_raw(kind, term, k), term, k)
            f.set_status(kind, term, 1)


def assemble_or_load(ctx: lib.Context, f: IntegralFile, kinds=("ol", "ke", "np", "ep")):
    """The resume of the integral stage: per kind (per potential term for
    EP) the matrices come from the file when its group's status is 1, else
    the device assembles them and they are written and marked.  Returns
    {kind: "loaded" | "assembled"}."""
    how = {}
    for kind in kinds:
        done = all(f.status(kind, t) == 1 for t in _terms(f, kind))
        if done:
            for term in _terms(f, kind):
                for k in range(f.num_kpoints):
                    ctx.set_packed(kind, f.read(kind, term, k), term, k)
            how[kind] = "loaded"
        else:
            ctx.assemble(kind)
            save_integrals(ctx, f, (kind,))
            how[kind] = "assembled"
    return how
