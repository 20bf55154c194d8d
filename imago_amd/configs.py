"""Seeded synthetic workloads of the integral + eigensolve path.

* ``si2_primitive``   diamond Si, 2-atom FCC primitive cell, a = 5.43 A,
                      1s2s2p3s3p (9 functions/atom), 12 exponents 0.12..2e4,
                      16 potential exponents 0.1..1e4, 4x4x4 MP mesh
                      (10 irreducible k), lattice cutoff 14 bohr.
* ``si64_supercell``  2x2x2 conventional Si (64 atoms), 0.05 A seeded
                      displacements, +3d (14 functions/atom, n = 896),
                      2x2x2 MP mesh (4 irreducible k), cutoff 16 bohr,
                      one potential type per site (potDim = 1024).
* ``sio2_amorphous``  216-atom a-SiO2 (72 Si, 144 O) by random sequential
                      insertion at 2.2 g/cm^3, s,p,d on both elements
                      (n = 1944), Gamma only, potDim = 3456.

Inputs are synthetic (no basis or potential database is shipped): radial
coefficients are seeded N(0,1) and normalised on site, potential
coefficients are seeded.  A "lattice cutoff" R is mapped onto the
reference's own screening knob: the most diffuse primitive pair is
screened at distance R, i.e. -ln(basisFnConvg) = R^2 * alpha_min / 2
(alphaDist = L (a1+a2)/(a1 a2), gaussRelations.f90:123-129).
"""
from __future__ import annotations

import math

import numpy as np

from .system import (AtomType, PotType, System, BOHR_PER_ANGSTROM,
                     geometric_alphas, onsite_radial_norm, kpoint_mesh,
                     cubic_point_group)

SI, O, C = 0, 1, 2
Z_OF = {SI: 14.0, O: 8.0, C: 6.0}
NUC_ALPHA = {SI: 20.0, O: 20.0, C: 20.0}     # atomicPDB si/pot1 nucAlpha
COV_RADIUS = {SI: 2.10, O: 1.25, C: 1.45}    # covalent radii in bohr (1.11, 0.66, 0.77 A)


def log_thresh_for_cutoff(cutoff_bohr: float, alpha_min: float) -> float:
    return cutoff_bohr ** 2 * alpha_min / 2.0


def random_shells(rng, alphas, ls):
    shells = []
    for l in ls:
        c = rng.normal(size=len(alphas))
        c = c / math.sqrt(onsite_radial_norm(alphas, c, l))
        shells.append((l, c))
    return shells


def make_atom_type(rng, element, ls, nalpha=12, amin=0.12, amax=2.0e4):
    alphas = geometric_alphas(amin, amax, nalpha)
    return AtomType(element=element, alphas=alphas,
                    shells=random_shells(rng, alphas, ls))


def diamond_conventional(a_bohr):
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    frac = np.concatenate([fcc, fcc + 0.25])
    return frac * a_bohr


def pot_coefficients(system: System, seed: int = 7) -> np.ndarray:
    """Seeded potential coefficients c_j (one spin), a screened-atom-like
    attractive profile with noise."""
    rng = np.random.default_rng(seed)
    out = []
    for t, a in system.pot_terms():
        alpha = system.pot_types[t].alphas[a]
        out.append(-0.4 * math.exp(-0.15 * math.log(alpha)) * (1.0 + 0.3 * rng.normal()))
    return np.array(out)


def si2_primitive(seed: int = 1, cutoff_bohr: float = 14.0, mesh=(4, 4, 4),
                  log_thresh: float | None = None, nalpha: int = 12,
                  npot: int = 16, ls=(0, 0, 1, 0, 1), core: bool = False) -> System:
    """Si2 primitive cell, 1s 2s 2p 3s 3p per atom: all valence (n = 18),
    or with core=True the first three shells as core states (core 10 +
    valence 8)."""
    rng = np.random.default_rng(seed)
    a = 5.43 * BOHR_PER_ANGSTROM
    rv = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.float64).T
    pos = np.array([[0, 0, 0], [0.25 * a] * 3])
    at = make_atom_type(rng, SI, ls, nalpha=nalpha)
    if core:
        at.core_shells, at.shells = at.shells[:3], at.shells[3:]
    pot = PotType(element=SI, alphas=geometric_alphas(0.1, 1.0e4, npot),
                  nuc_charge=Z_OF[SI], nuc_alpha=NUC_ALPHA[SI], sites=[0, 1],
                  covalent_radius=COV_RADIUS[SI])
    if log_thresh is None:
        log_thresh = log_thresh_for_cutoff(cutoff_bohr, at.alphas[0])
    kp, kw = kpoint_mesh(rv, mesh, ops=cubic_point_group())
    return System(real_vectors=rv, atom_types=[at], atom_type_of=[0, 0],
                  atom_pos=pos, pot_types=[pot], pot_site_pos=pos.copy(),
                  kpoints=kp, kpoint_weights=kw, log_basis_fn_thresh=log_thresh,
                  name="si2_primitive_core" if core else "si2_primitive")


def si64_supercell(seed: int = 2, cutoff_bohr: float = 16.0, mesh=(2, 2, 2),
                   sigma_angstrom: float = 0.05, log_thresh: float | None = None,
                   reps=(2, 2, 2), ls=(0, 0, 1, 0, 1, 2), gamma=False) -> System:
    rng = np.random.default_rng(seed)
    a = 5.43 * BOHR_PER_ANGSTROM
    conv = diamond_conventional(a)
    pos = []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                pos.append(conv + a * np.array([i, j, k]))
    pos = np.concatenate(pos)
    pos = pos + rng.normal(scale=sigma_angstrom * BOHR_PER_ANGSTROM, size=pos.shape)
    rv = np.diag([a * reps[0], a * reps[1], a * reps[2]]).astype(np.float64)
    at = make_atom_type(rng, SI, ls)
    palphas = geometric_alphas(0.1, 1.0e4, 16)
    pots = [PotType(element=SI, alphas=palphas, nuc_charge=Z_OF[SI],
                    nuc_alpha=NUC_ALPHA[SI], sites=[i], covalent_radius=COV_RADIUS[SI])
            for i in range(len(pos))]
    if log_thresh is None:
        log_thresh = log_thresh_for_cutoff(cutoff_bohr, at.alphas[0])
    kp, kw = kpoint_mesh(rv, mesh)          # no symmetry beyond time reversal
    return System(real_vectors=rv, atom_types=[at],
                  atom_type_of=np.zeros(len(pos), np.int32), atom_pos=pos,
                  pot_types=pots, pot_site_pos=pos.copy(), kpoints=kp,
                  kpoint_weights=kw, log_basis_fn_thresh=log_thresh, gamma=gamma,
                  name="si%d_supercell" % len(pos))


def _rsa(rng, box, n_si, n_o, d_sio, d_same, max_tries=200000):
    species, pos = [], []

    def ok(p, sp):
        for q, sq in zip(pos, species):
            d = p - q
            d -= box * np.round(d / box)
            r = math.sqrt(float(d @ d))
            if r < (d_same if sq == sp else d_sio):
                return False
        return True

    for sp, count in ((SI, n_si), (O, n_o)):
        placed = 0
        tries = 0
        while placed < count:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("random sequential insertion did not converge")
            p = rng.uniform(0, box, size=3)
            if ok(p, sp):
                pos.append(p)
                species.append(sp)
                placed += 1
    return np.array(pos), np.array(species)


def sio2_amorphous(seed: int = 3, n_si: int = 72, n_o: int = 144,
                   density_g_cm3: float = 2.2, cutoff_bohr: float | None = 16.0,
                   log_thresh: float | None = None) -> System:
    """Amorphous SiO2 by random sequential insertion (1.45 A Si-O, 2.3 A
    O-O / Si-Si minimum distances) in a cube at the given density; Si s,p,d
    with 12 primitives, O s,p,d with 10; one potential type per site;
    Gamma only.  The lattice cutoff is mapped onto -ln(basisFnConvg) with
    the most diffuse exponent (0.12), as for the Si cells; cutoff_bohr=None
    keeps the reference default 1e-16."""
    rng = np.random.default_rng(seed)
    mass = n_si * 28.0855 + n_o * 15.999
    vol_a3 = mass / 6.02214076e23 / density_g_cm3 * 1e24
    box_a = vol_a3 ** (1.0 / 3.0)
    pos_a, species = _rsa(rng, box_a, n_si, n_o, 1.45, 2.3)
    pos = pos_a * BOHR_PER_ANGSTROM
    box = box_a * BOHR_PER_ANGSTROM
    t_si = make_atom_type(rng, SI, (0, 1, 2), nalpha=12, amin=0.12, amax=2.0e4)
    t_o = make_atom_type(rng, O, (0, 1, 2), nalpha=10, amin=0.15, amax=1.0e4)
    types = np.where(species == SI, 0, 1).astype(np.int32)
    pots = []
    for i, sp in enumerate(species):
        pots.append(PotType(element=int(sp), alphas=geometric_alphas(0.1, 1.0e4, 16),
                            nuc_charge=Z_OF[int(sp)], nuc_alpha=NUC_ALPHA[int(sp)],
                            sites=[i], covalent_radius=COV_RADIUS[int(sp)]))
    rv = np.diag([box, box, box])
    if log_thresh is None:
        log_thresh = (log_thresh_for_cutoff(cutoff_bohr, 0.12) if cutoff_bohr is not None
                      else -math.log(1e-16))
    return System(real_vectors=rv, atom_types=[t_si, t_o], atom_type_of=types,
                  atom_pos=pos, pot_types=pots, pot_site_pos=pos.copy(),
                  kpoints=np.zeros((1, 3)), kpoint_weights=np.array([2.0]),
                  log_basis_fn_thresh=log_thresh, gamma=True, name="sio2_amorphous")


def small_test_system(seed: int = 5, gamma: bool = False, lmax: int = 2,
                      nalpha: int = 5, cutoff_bohr: float = 9.0,
                      two_elements: bool = True, with_f: bool = False,
                      core: bool = False) -> System:
    """A few-atom cell with mixed elements and a lopsided basis (different
    numOrbAlphas per l) that the CPU oracle assembles in seconds; core=True
    gives the first type s and p core shells and the second an s one."""
    rng = np.random.default_rng(seed)
    rv = np.array([[7.1, 0.3, 0.0], [0.0, 6.6, 0.4], [0.5, 0.0, 7.4]]).T
    pos = np.array([[0.1, 0.2, 0.0], [3.2, 2.9, 3.1], [1.4, 4.8, 2.2]])
    ls_a = [0, 1] + ([2] if lmax >= 2 else []) + ([3] if with_f else [])
    alphas = geometric_alphas(0.2, 60.0, nalpha)
    noa = [nalpha, nalpha - 1 if nalpha > 2 else nalpha, max(nalpha - 2, 1) if lmax >= 2 else 0,
           max(nalpha - 3, 1) if with_f else 0]
    ta = AtomType(element=SI, alphas=alphas, shells=[], num_orb_alphas=tuple(noa))
    ta.shells = random_shells(rng, alphas, ls_a)
    ta.shells = [(l, np.where(np.arange(nalpha) < noa[l], c, 0.0)) for l, c in ta.shells]
    types = [ta]
    tof = [0, 0, 0]
    if two_elements:
        alphas_b = geometric_alphas(0.3, 40.0, nalpha - 1)
        tb = AtomType(element=O, alphas=alphas_b,
                      shells=random_shells(rng, alphas_b, [0, 1]))
        types.append(tb)
        tof = [0, 1, 0]
    if core:
        crng = np.random.default_rng(seed + 1000)
        ta.core_shells = [(l, np.where(np.arange(nalpha) < noa[l], c, 0.0))
                          for l, c in random_shells(crng, alphas, [0, 1])]
        if two_elements:
            types[1].core_shells = random_shells(crng, alphas_b, [0])
    pots = [PotType(element=SI, alphas=geometric_alphas(0.15, 200.0, 4), nuc_charge=4.0,
                    nuc_alpha=5.0, sites=[0, 1] if not two_elements else [0],
                    covalent_radius=COV_RADIUS[SI]),
            ]
    if two_elements:
        pots.append(PotType(element=O, alphas=geometric_alphas(0.2, 100.0, 3),
                            nuc_charge=6.0, nuc_alpha=7.0, sites=[1],
                            covalent_radius=COV_RADIUS[O]))
        pots.append(PotType(element=SI, alphas=geometric_alphas(0.15, 200.0, 4),
                            nuc_charge=4.0, nuc_alpha=5.0, sites=[2],
                            covalent_radius=COV_RADIUS[SI]))
    else:
        pots.append(PotType(element=SI, alphas=geometric_alphas(0.15, 200.0, 4),
                            nuc_charge=4.0, nuc_alpha=5.0, sites=[2],
                            covalent_radius=COV_RADIUS[SI]))
    kp, kw = kpoint_mesh(rv, (2, 2, 1), shift=(0.5, 0.0, 0.0))
    lt = log_thresh_for_cutoff(cutoff_bohr, 0.2)
    return System(real_vectors=rv, atom_types=types, atom_type_of=tof,
                  atom_pos=pos, pot_types=pots, pot_site_pos=pos.copy(),
                  kpoints=kp, kpoint_weights=kw, log_basis_fn_thresh=lt,
                  gamma=gamma, name="small")
