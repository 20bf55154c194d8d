"""CPU restatement of the reference's density of states: test
infrastructure only (the checker of imago_dos and of imago_amd/pscf.py;
nothing in the product imports it).

* ``pdos_index``    pdosIndex and cumulDOSTotal (src/imago/dos.F90:353-560)
                    for detailCodePDOS 0-3, written loop by loop as the
                    reference's nested loops over atoms, l and radial
                    functions
* ``compute_dos``   computeDOS's Gaussian path (dos.F90:595-846): per k-point
                    and state the Mulliken accumulation of waveFnSqrd
                    against the overlap column (640-679), the electron
                    numbers of states at or below zero, pdosAccum and the
                    localisation index (680-700), the broadening with the
                    exp(-x^2) < 50 cut (707-743), then the electron factor,
                    the total DOS, the states in range and the integrated
                    area (756-846)
* ``path_kpoints``  makePathKPoints (src/imago/kpoints.f90:611-834)

Energies in Hartree as in the reference; eigenvectors C[k] are n x ns
(columns are states), the overlap S[k] full Hermitian n x n.
"""
from __future__ import annotations

import math

import numpy as np

HARTREE = 27.211386245988   # constants.f90:37


def pdos_index(atom_type_of, radial_l_lists, states_per_type, detail):
    """radial_l_lists[t]: the l of every valence radial function of type t
    (valeQN_lList); returns (pdosIndex 0-based per basis function,
    cumulDOSTotal)."""
    ntypes = len(radial_l_lists)
    num_ql = [[sum(1 for l in ls if l == q) for q in range(4)] for ls in radial_l_lists]
    if detail == 0:
        cumul = [0]
        for t in range(ntypes):
            cumul.append(cumul[-1] + sum(num_ql[t]))
        total = cumul[-1]
    elif detail == 1:
        total = len(atom_type_of)
    elif detail == 2:
        cumul = [0]
        for a, t in enumerate(atom_type_of):
            cumul.append(cumul[-1] + sum(num_ql[t]))
        total = cumul[-1]
    else:
        cumul = [0]
        for a, t in enumerate(atom_type_of):
            q = num_ql[t]
            cumul.append(cumul[-1] + q[0] + 3 * q[1] + 5 * q[2] + 7 * q[3])
        total = cumul[-1]
    idx = []
    vale = 0
    for a, t in enumerate(atom_type_of):
        if detail in (0, 2):
            q = num_ql[t]
            c0 = cumul[t] if detail == 0 else cumul[a]
            init = [c0, c0 + q[0], c0 + q[0] + q[1], c0 + q[0] + q[1] + q[2]]
            for j in range(q[0]):
                idx.append(init[0] + j)
            for j in range(q[1]):
                idx += [init[1] + j] * 3
            for j in range(q[2]):
                idx += [init[2] + j] * 5
            for j in range(q[3]):
                idx += [init[3] + j] * 7
        elif detail == 1:
            idx += [a] * states_per_type[t]
        else:
            for _ in range(states_per_type[t]):
                idx.append(vale)
                vale += 1
    return np.array(idx, dtype=np.int64), total


def compute_dos(E, C, S, kweights, spin, pdos_idx, ntot, emin, emax, delta, sigma,
                populations_total):
    """computeDOS for one spin (the Gaussian path).  E[k][j] shifted
    eigenvalues, C[k] (n x ns), S[k] (n x n); the grid has
    int((emax - emin) / delta) points from emin.  Returns a dict with
    pdosComplete before and after the electron factor, electronNumber per
    basis function, localizationIndex, totalSystemDos, the factor, the
    states in range and the integrated area."""
    E = np.asarray(E, dtype=np.float64)
    nk, ns = E.shape
    n = C[0].shape[0]
    sigma_sqrt_pi = math.sqrt(math.pi) * sigma
    npts = int((emax - emin) / delta)
    scale = emin + np.arange(npts) * delta
    pdos = np.zeros((ntot, npts))
    electron_number = np.zeros(n)
    loc = np.zeros(ns)
    for k in range(nk):
        Ck, Sk = np.asarray(C[k]), np.asarray(S[k])
        w = kweights[k]
        for j in range(ns):
            occupancy = 1.0 / spin if E[k, j] <= 0.0 else 0.0
            acc = np.zeros(ntot)
            c = Ck[:, j]
            for v in range(n):
                wave = np.conj(c[v]) * c
                one = float(np.sum(wave.real * Sk[:, v].real + wave.imag * Sk[:, v].imag))
                electron_number[v] += one * w * occupancy
                acc[pdos_idx[v]] += one / spin
                loc[j] += one * one * w / spin
            x2 = ((E[k, j] - scale) / sigma) ** 2
            on = x2 < 50.0
            factor = np.where(on, np.exp(-np.where(on, x2, 0.0)) / sigma_sqrt_pi / HARTREE * w, 0.0)
            pdos += np.outer(acc, factor)
    tdos_raw = pdos.sum(axis=0)
    states_in_range = sum(kweights[k] / spin for k in range(nk) for j in range(ns)
                          if emin < E[k, j] < emax)
    computed = electron_number.sum()
    efactor = populations_total / computed
    tdos = tdos_raw * efactor
    area = float(np.sum(tdos[:-1] + tdos[1:]) * delta * 0.5 * HARTREE)
    return dict(pdos_raw=pdos, pdos=pdos * efactor, electron_number=electron_number * efactor,
                electron_number_raw=electron_number, localization=loc, tdos=tdos,
                factor=efactor, states_in_range=states_in_range, area=area)


def path_kpoints(recip, paths, num_points):
    """makePathKPoints with the high-symmetry points in fractional
    coordinates of the reciprocal vectors (columns of recip)."""
    hs = [[recip @ np.asarray(p, dtype=np.float64) for p in path] for path in paths]
    dist_vec, dist_mag = [], []
    for i, path in enumerate(hs):
        dv, dm = [], []
        for j, p in enumerate(path):
            if i == 0 and j == 0:
                dv.append(np.zeros(3))
                dm.append(0.0)
            elif j == 0:
                dv.append(dist_vec[i - 1][-1])
                dm.append(dist_mag[i - 1][-1])
            else:
                d = p - path[j - 1]
                dv.append(d)
                dm.append(dm[j - 1] + math.sqrt(float(np.dot(d, d))))
        dist_vec.append(dv)
        dist_mag.append(dm)
    kpts, mags = [], []
    if num_points == sum(len(p) for p in hs):
        for i, path in enumerate(hs):
            for j, p in enumerate(path):
                kpts.append(p)
                mags.append(dist_mag[i][j])
    else:
        average = dist_mag[-1][-1] / (num_points - 1)
        delta = np.zeros(3)
        for i, path in enumerate(hs):
            for j, p in enumerate(path):
                if j != 0:
                    nseg = int((dist_mag[i][j] - dist_mag[i][j - 1]) / average)
                    if len(kpts) + nseg > num_points:
                        nseg = num_points - len(kpts)
                    if nseg >= 1:
                        delta = dist_vec[i][j] / nseg
                        for _ in range(1, nseg):
                            kpts.append(kpts[-1] + delta)
                            mags.append(mags[-1] + math.sqrt(float(np.dot(delta, delta))))
                kpts.append(p)
                if i == 0 and j == 0:
                    mags.append(0.0)
                elif j == 0:
                    mags.append(mags[-1])
                else:
                    mags.append(mags[-1] + math.sqrt(float(np.dot(delta, delta))))
    return np.array(kpts), np.array(mags), np.full(len(kpts), 2.0 / len(kpts))
