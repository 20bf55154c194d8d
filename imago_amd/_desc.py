"""ctypes mirrors of ``imago_system_desc`` and ``imago_xc_mesh_params``
(include/imago_hip.h)."""
from __future__ import annotations

import ctypes

import numpy as np

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class SystemDesc(ctypes.Structure):
    _fields_ = [
        ("num_atom_types", ctypes.c_int),
        ("max_alphas", ctypes.c_int),
        ("max_states", ctypes.c_int),
        ("type_element", _ip),
        ("type_num_orb_alphas", _ip),
        ("type_alphas", _dp),
        ("type_num_states", _ip),
        ("type_basis", _dp),
        ("type_lm_index", _ip),
        ("num_atoms", ctypes.c_int),
        ("atom_type", _ip),
        ("atom_pos", _dp),
        ("num_pot_types", ctypes.c_int),
        ("max_pot_alphas", ctypes.c_int),
        ("pot_type_element", _ip),
        ("pot_type_num_alphas", _ip),
        ("pot_type_alphas", _dp),
        ("pot_type_nuc_charge", _dp),
        ("pot_type_nuc_alpha", _dp),
        ("pot_type_first_site", _ip),
        ("pot_type_multiplicity", _ip),
        ("num_pot_sites", ctypes.c_int),
        ("pot_site_type", _ip),
        ("pot_site_pos", _dp),
        ("real_vectors", ctypes.c_double * 9),
        ("num_cells", ctypes.c_int),
        ("cell_dims", _dp),
        ("cell_sizes", _dp),
        ("num_kpoints", ctypes.c_int),
        ("kpoints", _dp),
        ("kpoint_weights", _dp),
        ("gamma", ctypes.c_int),
        ("log_basis_fn_thresh", ctypes.c_double),
        ("type_num_core_states", _ip),
    ]


class XcMeshParams(ctypes.Structure):
    _fields_ = [
        ("num_sample_vectors", ctypes.c_int),
        ("weight_in", ctypes.c_double),
        ("weight_out", ctypes.c_double),
        ("r_sample_in", ctypes.c_double),
        ("r_sample_out", ctypes.c_double),
        ("r_sample_space", ctypes.c_double),
    ]


def make_xc_params(mesh):
    """XcMeshParams from an XcMesh (imago_amd.system)."""
    return XcMeshParams(int(mesh.num_sample_vectors), float(mesh.weight_in),
                        float(mesh.weight_out), float(mesh.r_sample_in),
                        float(mesh.r_sample_out), float(mesh.r_sample_space))


def make_desc(system):
    """Build a SystemDesc for ``system``; returns (desc, keepalive list)."""
    arr = system.desc_arrays()
    keep = []

    def d(name):
        a = np.ascontiguousarray(arr[name], dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(_dp)

    def i(name):
        a = np.ascontiguousarray(arr[name], dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(_ip)

    s = SystemDesc()
    s.num_atom_types = len(arr["type_element"])
    s.max_alphas = arr["max_alphas"]
    s.max_states = arr["max_states"]
    s.type_element = i("type_element")
    s.type_num_orb_alphas = i("type_num_orb_alphas")
    s.type_alphas = d("type_alphas")
    s.type_num_states = i("type_num_states")
    s.type_basis = d("type_basis")
    s.type_lm_index = i("type_lm_index")
    s.num_atoms = len(arr["atom_type"])
    s.atom_type = i("atom_type")
    s.atom_pos = d("atom_pos")
    s.num_pot_types = len(arr["pot_type_element"])
    s.max_pot_alphas = arr["max_pot_alphas"]
    s.pot_type_element = i("pot_type_element")
    s.pot_type_num_alphas = i("pot_type_num_alphas")
    s.pot_type_alphas = d("pot_type_alphas")
    s.pot_type_nuc_charge = d("pot_type_nuc_charge")
    s.pot_type_nuc_alpha = d("pot_type_nuc_alpha")
    s.pot_type_first_site = i("pot_type_first_site")
    s.pot_type_multiplicity = i("pot_type_multiplicity")
    s.num_pot_sites = len(arr["pot_site_type"])
    s.pot_site_type = i("pot_site_type")
    s.pot_site_pos = d("pot_site_pos")
    for q in range(9):
        s.real_vectors[q] = float(arr["real_vectors"][q])
    s.num_cells = len(arr["cell_sizes"])
    s.cell_dims = d("cell_dims")
    s.cell_sizes = d("cell_sizes")
    s.num_kpoints = len(arr["kpoint_weights"])
    s.kpoints = d("kpoints")
    s.kpoint_weights = d("kpoint_weights")
    s.gamma = arr["gamma"]
    s.log_basis_fn_thresh = arr["log_basis_fn_thresh"]
    if np.any(arr["type_num_core_states"]):
        s.type_num_core_states = i("type_num_core_states")
    return s, keep
