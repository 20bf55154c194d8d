"""Post-SCF band structure and density of states on the device.

With the converged potential coefficients of an SCF, the reference's
post-SCF runs (src/imago/imago.F90:740-945) assemble the Hamiltonian at new
k-points with the potential folded into one accumulation
(integrals.F90:2860-3605), solve it (secularEqnPSCF, secularEqn.F90:
331-617), find the highest occupied energy (populateStates), shift the
eigenvalues by it (shiftEnergyEigenValues, secularEqn.F90:846-868) and write

  * the band structure along a path of high-symmetry points
    (makePathKPoints, kpoints.f90:611-834; printSYBD, imago.F90:800-888,
    fort.31), and
  * the Gaussian-broadened total and partial density of states with the
    Mulliken populations (computeDOS, dos.F90:160-1231; fort.60 TDOS,
    fort.70 PDOS, fort.80 localisation index).

Here the integrals are the library's direct EP path (no potDim panels, any
number of k-points), the eigenpairs come from imago_secular and the
populations, channel sums and broadening from imago_dos.  The host keeps
what the reference keeps on the host: the path, the channel map
(pdosIndex), the electron-count normalisation and the output files.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import lib, scf
from .system import System, reciprocal_vectors

HARTREE_EV = 27.211386245988   # constants.f90:37


def path_kpoints(real_vectors, paths, num_points, cartesian=False):
    """makePathKPoints (kpoints.f90:611-834): the k-points of a band path.

    paths: list of paths, each a list of high-symmetry points (fractional
    coordinates of the reciprocal vectors unless cartesian); num_points the
    requested count (NUM_PATH_KP).  Returns (kpoints (nk, 3) Cartesian,
    path magnitudes (nk,), weights (nk,) = 2 / nk, indices of the
    high-symmetry points)."""
    recip, _ = reciprocal_vectors(np.asarray(real_vectors, dtype=np.float64))
    hs = []
    for p in paths:
        pts = np.asarray(p, dtype=np.float64).reshape(-1, 3)
        hs.append(pts.copy() if cartesian else pts @ recip.T)
    # cumulative distances along the paths (a new path starts at the
    # distance the previous one reached)
    dvec, dmag = [], []
    for i, pts in enumerate(hs):
        dv = np.zeros_like(pts)
        dm = np.zeros(len(pts))
        for j in range(len(pts)):
            if i == 0 and j == 0:
                dv[j], dm[j] = 0.0, 0.0
            elif j == 0:
                dv[j], dm[j] = dvec[i - 1][-1], dmag[i - 1][-1]
            else:
                dv[j] = pts[j] - pts[j - 1]
                dm[j] = dm[j - 1] + math.sqrt(float(dv[j] @ dv[j]))
        dvec.append(dv)
        dmag.append(dm)
    total_hs = sum(len(p) for p in hs)
    kp, mag, marks = [], [], []
    if num_points == total_hs:
        for i, pts in enumerate(hs):
            for j in range(len(pts)):
                kp.append(pts[j].copy())
                mag.append(dmag[i][j])
                marks.append(len(kp) - 1)
    else:
        avg = dmag[-1][-1] / (num_points - 1)
        seg = np.zeros(3)
        for i, pts in enumerate(hs):
            for j in range(len(pts)):
                if j != 0:
                    nseg = int((dmag[i][j] - dmag[i][j - 1]) / avg)
                    if len(kp) + nseg > num_points:
                        nseg = num_points - len(kp)
                    if nseg >= 1:
                        seg = dvec[i][j] / nseg
                        for _ in range(nseg - 1):
                            kp.append(kp[-1] + seg)
                            mag.append(mag[-1] + math.sqrt(float(seg @ seg)))
                kp.append(pts[j].copy())
                marks.append(len(kp) - 1)
                if i == 0 and j == 0:
                    mag.append(0.0)
                elif j == 0:
                    mag.append(mag[-1])
                else:
                    mag.append(mag[-1] + math.sqrt(float(seg @ seg)))
    nk = len(kp)
    return np.array(kp), np.array(mag), np.full(nk, 2.0 / nk), marks


def with_kpoints(system: System, kpoints, weights) -> System:
    """The system with another k-point set (complex Bloch sums)."""
    return replace(system, kpoints=np.asarray(kpoints, dtype=np.float64).reshape(-1, 3),
                   kpoint_weights=np.asarray(weights, dtype=np.float64), gamma=False,
                   name=system.name + "_pscf")


def radial_counts(atom_type):
    """numQN_lValeRadialFns: valence radial functions per l (s, p, d, f)."""
    c = [0, 0, 0, 0]
    for l, _ in atom_type.shells:
        c[l] += 1
    return c


def pdos_channels(system: System, detail: int):
    """pdosIndex and cumulDOSTotal of computeDOS (dos.F90:353-560): the PDOS
    channel (0-based) of every basis function for detailCodePDOS 0 (every
    radial function of every type), 1 (every atom), 2 (every radial
    function of every atom) or 3 (every basis function).  Within an atom
    the reference assigns the states in the order of its counts per l (all
    s radial functions, then three states for each p, five for each d,
    seven for each f), as this does."""
    types = system.atom_types
    counts = [radial_counts(t) for t in types]
    per_type = [sum(c) for c in counts]
    cumul_type = np.concatenate([[0], np.cumsum(per_type)]).astype(int)
    chan = []
    if detail == 1:
        for a, t in enumerate(system.atom_type_of):
            chan += [a] * types[t].num_states
        return np.array(chan, dtype=np.int32), system.num_atoms
    if detail == 3:
        return np.arange(system.vale_dim, dtype=np.int32), system.vale_dim
    if detail not in (0, 2):
        raise ValueError("detailCodePDOS is 0, 1, 2 or 3")
    off = 0
    for a, t in enumerate(system.atom_type_of):
        base = cumul_type[t] if detail == 0 else off
        c = counts[t]
        first = [base, base + c[0], base + c[0] + c[1], base + c[0] + c[1] + c[2]]
        for l in range(4):
            for j in range(c[l]):
                chan += [first[l] + j] * (2 * l + 1)
        off += per_type[t]
    total = int(cumul_type[-1]) if detail == 0 else off
    return np.array(chan, dtype=np.int32), total


@dataclass
class DosResult:
    energy_ev: np.ndarray            # the energy scale, eV
    tdos: np.ndarray                 # states / eV / cell, after the electron factor
    pdos: np.ndarray                 # [channel, point], after the electron factor
    electron_number: np.ndarray      # per basis function, after the electron factor
    localization: np.ndarray         # per state
    electron_factor: float
    electrons_computed: float
    electrons_expected: float
    integrated_area: float
    states_in_range: float
    occupied_energy: float           # Hartree, the shift applied
    eigenvalues: np.ndarray          # shifted, Hartree [nk, ns]
    channel: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))


def density_of_states(ctx, system: System, coeffs, sigma_ev=0.2, emin_ev=-30.0,
                      emax_ev=30.0, delta_ev=0.05, detail=1, num_states=None,
                      electrons=None, spin=1, setup=True) -> DosResult:
    """dos (imago.F90:890-945) with computeDOS's Gaussian path on the
    device.  The system's k-points are the integration mesh; coeffs are
    the converged potential coefficients.  Energies in eV at the boundary,
    Hartree inside as in the reference (input.f90 converts)."""
    if setup:
        ctx.set_system(system)
        for kind in ("ol", "ke", "np", "ep_direct"):
            ctx.assemble(kind)
    n = system.vale_dim
    ns = n if num_states is None else num_states
    ev = ctx.secular(np.asarray(coeffs, dtype=np.float64), ns)
    nel = scf.num_electrons(system) if electrons is None else electrons
    occ, (e_occ, _, _) = lib.populate(ev, system.kpoint_weights, nel)
    shifted = ev - e_occ
    sigma, emin, emax, delta = (x / HARTREE_EV for x in (sigma_ev, emin_ev, emax_ev, delta_ev))
    npts = int((emax - emin) / delta)
    channel, nch = pdos_channels(system, detail)
    pd, en, loc = ctx.dos(shifted, spin, channel, nch, emin, delta, npts, sigma)
    tdos = pd.sum(axis=0)
    w = np.asarray(system.kpoint_weights)
    states_in_range = float(sum(w[k] / spin for k in range(shifted.shape[0])
                                for e in shifted[k] if emin < e < emax))
    computed = float(en.sum())
    expected = float(occ.sum())
    factor = expected / computed
    tdos = tdos * factor
    pd = pd * factor
    en = en * factor
    area = float(np.sum(tdos[:-1] + tdos[1:]) * delta * 0.5 * HARTREE_EV)
    scale = (emin + np.arange(npts) * delta) * HARTREE_EV
    return DosResult(scale, tdos, pd, en, loc, factor, computed, expected, area,
                     states_in_range, e_occ, shifted, channel)


def write_tdos(path, res: DosResult):
    """fort.60 (dos.F90:864-871): header, then energy (eV) and TDOS."""
    with open(path, "w") as f:
        f.write("%14s%14s\n" % ("    ENERGY(eV)", "          TDOS"))
        for e, t in zip(res.energy_ev, res.tdos):
            f.write("%14.4f%14.6f\n" % (e, t))


def write_localization(path, res: DosResult, kpoint_weights):
    """fort.80 (dos.F90:1165-1188): per state the k-weighted energy and the
    localisation index, both halved (the k-point weights sum to two), the
    energy in eV."""
    w = np.asarray(kpoint_weights, dtype=np.float64)
    avg = (res.eigenvalues * w[:, None]).sum(axis=0) / 2.0 * HARTREE_EV
    with open(path, "w") as f:
        for e, l in zip(avg, res.localization / 2.0):
            f.write(" %24.8f  %24.8f\n" % (e, l))


@dataclass
class BandResult:
    kpoints: np.ndarray
    path_magnitude: np.ndarray
    high_symmetry: list
    eigenvalues: np.ndarray      # shifted, Hartree [nk, ns]
    occupied_energy: float


def band_structure(ctx, system: System, coeffs, paths, num_points, cartesian=False,
                   num_states=None, electrons=None) -> BandResult:
    """bandPSCF + printSYBD (imago.F90:740-888): the eigenvalues along the
    path, shifted by the highest occupied energy of the path's own
    population (weights 2 / nk)."""
    kp, mag, w, marks = path_kpoints(system.real_vectors, paths, num_points, cartesian)
    s = with_kpoints(system, kp, w)
    ctx.set_system(s)
    for kind in ("ol", "ke", "np", "ep_direct"):
        ctx.assemble(kind)
    ns = s.vale_dim if num_states is None else num_states
    ev = ctx.secular(np.asarray(coeffs, dtype=np.float64), ns)
    nel = scf.num_electrons(s) if electrons is None else electrons
    _, (e_occ, _, _) = lib.populate(ev, w, nel)
    return BandResult(kp, mag, marks, ev - e_occ, e_occ)


def write_bands(path, res: BandResult):
    """fort.31 (printSYBD, imago.F90:830-842): the k-point and state counts;
    per k-point its path magnitude on a line, then the shifted eigenvalues
    in eV, eleven to a line (format 11f14.5)."""
    nk, ns = res.eigenvalues.shape
    with open(path, "w") as f:
        f.write(" %11d %11d\n" % (nk, ns))
        for k in range(nk):
            f.write(" %24.16f\n" % res.path_magnitude[k])
            vals = res.eigenvalues[k] * HARTREE_EV
            for j0 in range(0, ns, 11):
                f.write("".join("%14.5f" % v for v in vals[j0:j0 + 11]) + "\n")
