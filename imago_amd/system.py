"""Host-side system setup: what the reference's Fortran host computes in
setupSCF before the integral stage (imago.F90:103-256) and hands to the
library through ``imago_set_system``.

* ``make_lattice``     lattice.f90:310-444 (initializeLattice),
                       getNumCells/makeLattice/cellBorderCheck (469-770) and
                       the merge sort of sortSubs.f90:18-177
* ``kpoint_mesh``      kpoints.f90:837-1224 (initializeKPointMesh: a
                       Monkhorst-Pack mesh folded by point-group operations
                       and time reversal) and convertKPointsToXYZ (543-576)
* ``AtomType``         basis.f90:55-243 (renormalizeBasis: lmIndex, the
                       angNorm-scaled basis coefficients)
* ``PotType``          potTypes.f90:170-195 (makePotAlphas)

The arrays produced here are plain numpy arrays laid out exactly as
``imago_system_desc`` in include/imago_hip.h expects.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903   # constants.f90 bohrRad

# basis.f90:91-106, the normalisation of the 16 real solid harmonics
ANG_NORM = np.array([
    0.2820947917738782, 0.4886025119029199, 0.4886025119029199,
    0.4886025119029199, 1.0925484305920790, 1.0925484305920790,
    1.0925484305920790, 0.5462742152960396, 0.3153915652525200,
    2.8906114426405543, 1.4453057213202771, 0.5900435899266435,
    0.5900435899266435, 0.3731763325901154, 0.4570457994644657,
    0.4570457994644657])


def geometric_alphas(amin: float, amax: float, n: int) -> np.ndarray:
    """makePotAlphas (potTypes.f90:170-195): a_i = amin*(amax/amin)^((i-1)/(n-1))."""
    if n == 1:
        return np.array([amin], dtype=np.float64)
    i = np.arange(n, dtype=np.float64)
    return amin * (amax / amin) ** (i / (n - 1))


def onsite_radial_norm(alphas: np.ndarray, coeffs: np.ndarray, l: int) -> float:
    """<phi|phi> of sum_a c_a Y_lm r^l exp(-a r^2) about one centre:
    sum_ab c_a c_b Gamma(l+3/2) / (2 (a+b)^(l+3/2))."""
    p = alphas[:, None] + alphas[None, :]
    g = math.gamma(l + 1.5) / 2.0
    return float(coeffs @ (g / p ** (l + 1.5)) @ coeffs)


@dataclass
class AtomType:
    """One atom type: an ascending exponent list shared by all shells and a
    list of contracted shells (l, radial coefficients over the first
    numOrbAlphas(l+1) exponents).  core_shells are the type's core
    functions, which the integral stage orthogonalises the valence
    functions against (intgOrtho.F90); the reference stores each type's
    valence states first, then its core states (basis.f90:160-243)."""
    element: int
    alphas: np.ndarray
    shells: list            # [(l, coeffs)] in valence order
    num_orb_alphas: tuple = (0, 0, 0, 0)
    core_shells: list = field(default_factory=list)

    def __post_init__(self):
        self.alphas = np.asarray(self.alphas, dtype=np.float64)
        if np.any(np.diff(self.alphas) <= 0):
            raise ValueError("atom-type exponents must be ascending")
        if self.num_orb_alphas == (0, 0, 0, 0):
            lmax = max(l for l, _ in list(self.shells) + list(self.core_shells))
            n = len(self.alphas)
            self.num_orb_alphas = tuple(n if l <= lmax else 0 for l in range(4))
        noa = self.num_orb_alphas
        if any(noa[l + 1] > noa[l] for l in range(3)):
            raise ValueError("numOrbAlphas must be non-increasing in l")

    @property
    def num_states(self) -> int:
        return sum(2 * l + 1 for l, _ in self.shells)

    @property
    def num_core_states(self) -> int:
        return sum(2 * l + 1 for l, _ in self.core_shells)

    def states(self):
        """(lmIndex 1..16, basis column over alphas) per valence state,
        basis.f90:160-243."""
        return self._states(self.shells)

    def core_states(self):
        """The same for the core states."""
        return self._states(self.core_shells)

    def _states(self, shells):
        out = []
        na = len(self.alphas)
        for l, c in shells:
            nal = self.num_orb_alphas[l]
            for k in range(1, 2 * l + 2):
                lm = l * l + k
                col = np.zeros(na)
                col[:nal] = np.asarray(c, dtype=np.float64)[:nal] * ANG_NORM[lm - 1]
                out.append((lm, col))
        return out


@dataclass
class XcMesh:
    """Exchange-correlation mesh input (readExchCorrMeshParameters,
    exchCorr.f90:876-913); defaults are makeinput's (num_samp_vectors 100,
    xc weights 0.5 0.5, radial sampling 0.1 3.5 0.8)."""
    num_sample_vectors: int = 100
    weight_in: float = 0.5
    weight_out: float = 0.5
    r_sample_in: float = 0.1
    r_sample_out: float = 3.5
    r_sample_space: float = 0.8


@dataclass
class PotType:
    element: int
    alphas: np.ndarray
    nuc_charge: float
    nuc_alpha: float
    sites: list             # indices into the potential-site list (contiguous)
    covalent_radius: float = 0.0   # bohr (potTypes(:)%covalentRadius)

    def __post_init__(self):
        self.alphas = np.asarray(self.alphas, dtype=np.float64)


# ---------------------------------------------------------------------------
# lattice.f90


def _merge_sort_order(data: np.ndarray) -> np.ndarray:
    """Order produced by mergeSort (sortSubs.f90:18-177) with one element per
    initial segment: bottom-up merges of adjacent runs in which, on equal
    keys, the element of the second run is taken first (the comparison is
    strict '<').  Two equal keys first meet when their runs merge, and the
    later one then goes first and stays first, so among equal keys the
    order is by descending position: a stable sort on (key, -position)."""
    data = np.asarray(data)
    return np.lexsort((-np.arange(len(data)), data)).astype(np.int64)


def _cell_border_hits(vectors, dims, negl) -> np.ndarray:
    """cellBorderCheck (lattice.f90:700-770) for every row of dims: the
    point is kept if one of the seven corner moves of the cell (by the sum
    of the lattice vectors, one axis at a time) comes within sqrt(negl) of
    the origin."""
    shift = vectors.sum(axis=1)
    t = np.array(dims, dtype=np.float64, copy=True)
    hit = np.zeros(len(t), dtype=bool)
    for ax, sg in [(0, +1), (1, +1), (0, -1), (2, +1), (0, +1), (1, -1), (0, -1)]:
        t[:, ax] += sg * shift[ax]
        hit |= (t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1] + t[:, 2] * t[:, 2]) < negl
    return hit


def negligible_limit_real(log_thresh: float, min_atomic_alpha: float,
                          min_pot_alpha: float, real_vectors: np.ndarray) -> float:
    """initializeLattice (lattice.f90:353-380)."""
    negl = log_thresh * max(2.0 / min_atomic_alpha,
                            1.0 / min_pot_alpha + 0.5 / min_atomic_alpha)
    max_sq = max(float(np.sum(real_vectors[:, i] ** 2)) for i in range(3))
    negl = negl + 0.75 * max_sq + math.sqrt(3.0 * max_sq * negl)
    return max(negl, 2.0 * log_thresh / min_atomic_alpha)


def reciprocal_vectors(real_vectors: np.ndarray):
    """recipVectors (columns) and realCellVolume as lattice.f90:195-216
    computes them: column i is 2 pi (a_{i+1} x a_{i+2}) / ((a_{i+1} x
    a_{i+2}) . a_i); the volume is the last of those triple products,
    made positive."""
    rv = np.asarray(real_vectors, dtype=np.float64)
    recip = np.zeros((3, 3))
    vol = 0.0
    for i in range(3):
        i1, i2 = (i + 1) % 3, (i + 2) % 3
        vol = 0.0
        for j in range(3):
            j1, j2 = (j + 1) % 3, (j + 2) % 3
            recip[j, i] = rv[j1, i1] * rv[j2, i2] - rv[j2, i1] * rv[j1, i2]
            vol = vol + recip[j, i] * rv[j, i]
        recip[:, i] = 2.0 * math.pi * recip[:, i] / vol
    return recip, abs(vol)


def make_lattice(real_vectors: np.ndarray, negl: float):
    """Sorted real-space cell list (makeLattice, lattice.f90:571-696).
    real_vectors: 3x3 with a,b,c as columns.  Returns (dims[n,3], sizes[n]).
    The candidate points are the reference's repetition box in its loop
    order (i, then j, then k), each position accumulated over the three
    vectors in the reference's order; a point is kept when |R|^2 <= negl
    or cellBorderCheck keeps it, then sorted as mergeSort sorts."""
    rv = np.asarray(real_vectors, dtype=np.float64)
    reps = [2 * math.ceil(math.sqrt(negl) / math.sqrt(np.sum(rv[:, i] ** 2))) + 1
            for i in range(3)]
    axes = [np.arange(1, reps[a] + 1) - (1 + reps[a]) // 2 for a in range(3)]
    r = [g.ravel() for g in np.meshgrid(*axes, indexing="ij")]
    d = np.zeros((len(r[0]), 3))
    for l in range(3):
        for m in range(3):
            d[:, m] = d[:, m] + rv[m, l] * r[l]
    sizes = d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2
    far = sizes > negl
    keep = ~far
    keep[far] = _cell_border_hits(rv, d[far], negl)
    dims, sizes = d[keep], sizes[keep]
    order = _merge_sort_order(sizes)
    return dims[order], sizes[order]


# ---------------------------------------------------------------------------
# kpoints.f90


def cubic_point_group() -> list:
    """The 48 operations of O_h as Cartesian signed permutation matrices."""
    ops = []
    import itertools
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for r in range(3):
                m[r, perm[r]] = signs[r]
            ops.append(m)
    return ops


def kpoint_mesh(real_vectors: np.ndarray, mesh, shift=(0.5, 0.5, 0.5),
                ops=None):
    """Monkhorst-Pack mesh folded to its irreducible wedge by the given
    Cartesian point operations plus time reversal (k ~ -k), as
    initializeKPointMesh does (kpoints.f90:837-1224).  Returns Cartesian k
    (2*pi included, kpoints.f90:543-576) and weights summing to 2
    (kpoints.f90:816)."""
    recip = 2.0 * math.pi * np.linalg.inv(real_vectors).T   # columns b1,b2,b3
    inv_recip = np.linalg.inv(recip)
    if ops is None:
        ops = [np.eye(3)]
    ops = list(ops) + [-o for o in ops]
    pts = []
    for i in range(mesh[0]):
        for j in range(mesh[1]):
            for k in range(mesh[2]):
                f = np.array([(i + shift[0]) / mesh[0], (j + shift[1]) / mesh[1],
                              (k + shift[2]) / mesh[2]])
                f = f - np.floor(f + 0.5)
                pts.append(f)
    pts = np.array(pts)

    def key(frac):
        g = frac - np.round(frac)
        g[np.abs(g - 0.5) < 1e-8] = -0.5
        return tuple(np.round(g, 8))

    seen = {}
    reps, weights = [], []
    for f in pts:
        kc = recip @ f
        kf = key(f)
        if kf in seen:
            weights[seen[kf]] += 1
            continue
        idx = len(reps)
        reps.append(kc)
        weights.append(1)
        for o in ops:
            seen.setdefault(key(inv_recip @ (o @ kc)), idx)
    w = np.array(weights, dtype=np.float64)
    return np.array(reps), 2.0 * w / w.sum()


# ---------------------------------------------------------------------------


@dataclass
class System:
    """A periodic system in the shape of ``imago_system_desc``."""
    real_vectors: np.ndarray            # 3x3, columns a, b, c (bohr)
    atom_types: list                    # [AtomType]
    atom_type_of: np.ndarray            # [natoms]
    atom_pos: np.ndarray                # [natoms, 3] bohr
    pot_types: list                     # [PotType]
    pot_site_pos: np.ndarray            # [npotsites, 3]
    kpoints: np.ndarray                 # [nk, 3] Cartesian
    kpoint_weights: np.ndarray          # [nk]
    log_basis_fn_thresh: float
    gamma: bool = False
    log_elec_thresh: float = -math.log(1.0e-16)   # -ln(electroConvg), lattice.f90:294
    cell_dims: np.ndarray = None
    cell_sizes: np.ndarray = None
    name: str = ""
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        self.atom_pos = np.ascontiguousarray(self.atom_pos, dtype=np.float64)
        self.pot_site_pos = np.ascontiguousarray(self.pot_site_pos, dtype=np.float64)
        self.atom_type_of = np.ascontiguousarray(self.atom_type_of, dtype=np.int32)
        if self.gamma:
            self.kpoints = np.zeros((1, 3))
            self.kpoint_weights = np.array([2.0])
        if self.cell_dims is None:
            min_a = min(float(t.alphas[0]) for t in self.atom_types)
            min_p = min(float(p.alphas[0]) for p in self.pot_types)
            negl = negligible_limit_real(self.log_basis_fn_thresh, min_a, min_p,
                                         self.real_vectors)
            self.cell_dims, self.cell_sizes = make_lattice(self.real_vectors, negl)

    # sizes -----------------------------------------------------------------
    @property
    def num_atoms(self) -> int:
        return len(self.atom_type_of)

    @property
    def vale_dim(self) -> int:
        return int(sum(self.atom_types[t].num_states for t in self.atom_type_of))

    @property
    def pot_dim(self) -> int:
        return int(sum(len(p.alphas) for p in self.pot_types))

    @property
    def core_dim(self) -> int:
        return int(sum(self.atom_types[t].num_core_states for t in self.atom_type_of))

    def all_valence(self) -> "System":
        """The system with every type's core shells appended to its valence
        shells: its matrices hold, per atom, the valence then the core
        functions (the extended blocks the orthogonalisation starts from)."""
        from dataclasses import replace
        types = [AtomType(element=t.element, alphas=t.alphas,
                          shells=list(t.shells) + list(t.core_shells),
                          num_orb_alphas=t.num_orb_alphas) for t in self.atom_types]
        return replace(self, atom_types=types, name=self.name + "_all_valence")

    @property
    def num_kpoints(self) -> int:
        return 1 if self.gamma else len(self.kpoints)

    def pot_terms(self):
        """(pot type, alpha index) per term j in the reference's order:
        types in the order of their first site, then alphas
        (integrals.F90:2411-2476)."""
        order = sorted(range(len(self.pot_types)),
                       key=lambda t: self.pot_types[t].sites[0])
        return [(t, a) for t in order for a in range(len(self.pot_types[t].alphas))]

    # electrostatics ----------------------------------------------------------
    @property
    def min_pot_alpha(self) -> float:
        """minPotAlpha (potTypes.f90:213-250)."""
        return min([10000.0] + [float(p.alphas[0]) for p in self.pot_types])

    def recip_cells(self, log_elec_thresh: float | None = None):
        """Reciprocal-space cells of initializeLattice (lattice.f90:386-420):
        |G|^2 <= 4 logElecThresh minPotAlpha, sorted by |G|^2."""
        th = self.log_elec_thresh if log_elec_thresh is None else log_elec_thresh
        recip, _ = reciprocal_vectors(self.real_vectors)
        return make_lattice(recip, 4.0 * th * self.min_pot_alpha)

    # packed arrays for imago_system_desc -------------------------------------
    def desc_arrays(self) -> dict:
        nt = len(self.atom_types)
        maxa = max(len(t.alphas) for t in self.atom_types)
        maxs = max(t.num_states + t.num_core_states for t in self.atom_types)
        type_alphas = np.zeros((nt, maxa))
        type_basis = np.zeros((nt, maxs, maxa))
        type_lm = np.zeros((nt, maxs), dtype=np.int32)
        noa = np.zeros((nt, 4), dtype=np.int32)
        nst = np.zeros(nt, dtype=np.int32)
        ncore = np.zeros(nt, dtype=np.int32)
        el = np.zeros(nt, dtype=np.int32)
        for ti, t in enumerate(self.atom_types):
            type_alphas[ti, :len(t.alphas)] = t.alphas
            noa[ti] = t.num_orb_alphas
            nst[ti] = t.num_states
            ncore[ti] = t.num_core_states
            el[ti] = t.element
            # valence states, then core states
            for si, (lm, col) in enumerate(t.states() + t.core_states()):
                type_lm[ti, si] = lm
                type_basis[ti, si, :len(col)] = col
        npt = len(self.pot_types)
        maxpa = max(len(p.alphas) for p in self.pot_types)
        pot_alphas = np.zeros((npt, maxpa))
        site_type = np.zeros(len(self.pot_site_pos), dtype=np.int32)
        for pi, p in enumerate(self.pot_types):
            pot_alphas[pi, :len(p.alphas)] = p.alphas
            if list(p.sites) != list(range(p.sites[0], p.sites[0] + len(p.sites))):
                raise ValueError("sites of a potential type must be contiguous")
            site_type[list(p.sites)] = pi
        return dict(
            type_element=el, type_num_orb_alphas=noa, type_alphas=type_alphas,
            type_num_states=nst, type_basis=type_basis, type_lm_index=type_lm,
            type_num_core_states=ncore,
            max_alphas=maxa, max_states=maxs,
            atom_type=self.atom_type_of, atom_pos=self.atom_pos,
            pot_type_element=np.array([p.element for p in self.pot_types], np.int32),
            pot_type_num_alphas=np.array([len(p.alphas) for p in self.pot_types], np.int32),
            pot_type_alphas=pot_alphas, max_pot_alphas=maxpa,
            pot_type_nuc_charge=np.array([p.nuc_charge for p in self.pot_types]),
            pot_type_nuc_alpha=np.array([p.nuc_alpha for p in self.pot_types]),
            pot_type_first_site=np.array([p.sites[0] for p in self.pot_types], np.int32),
            pot_type_multiplicity=np.array([len(p.sites) for p in self.pot_types], np.int32),
            pot_site_type=site_type, pot_site_pos=self.pot_site_pos,
            real_vectors=np.ascontiguousarray(self.real_vectors.T).ravel(),
            cell_dims=np.ascontiguousarray(self.cell_dims),
            cell_sizes=np.ascontiguousarray(self.cell_sizes),
            kpoints=np.ascontiguousarray(self.kpoints, dtype=np.float64),
            kpoint_weights=np.ascontiguousarray(self.kpoint_weights, dtype=np.float64),
            gamma=int(self.gamma), log_basis_fn_thresh=float(self.log_basis_fn_thresh),
        )
