"""CPU restatement of the reference's SCF loop: test infrastructure only.

This is the oracle's side of the self-consistent field iteration, the
checker of the library's SCF path and its CPU baseline (nothing in the
product imports it):

* ``run_scf``          mainSCF (src/imago/imago.F90:488-523): secular
                       equation, populateStates, makeValenceRho, makeSCFPot
                       until the potential converges
* ``ScfPotential``     makeSCFPot (src/imago/potentialUpdate.F90:53-1866)
                       for spin 1, no core states, non-relativistic, LDA
                       with the Wigner functional (xcCode 100): charge fit,
                       electrostatic potential and energy, exchange-
                       correlation potential and energy, total energy,
                       potential mixing and the RMS real-space convergence
                       test
* ``density_traces``   makeValenceRho's traces (src/imago/valeCharge.F90:
                       230-393): Tr(rho S/NP/KE) and potRho(j) = Tr(rho V_j)
                       per k-point through matrixElementMult
                       (matrixSubs.F90:66-112)

Every matrix the loop reads comes from the oracle's own restatement of the
integrals (oracle.assemble, orc_elecstat, orc_residual_q, the XC mesh of
orc_xc_*); solveDPOSVX is LAPACK dposvx('E', 'U') through SciPy, zhegv /
dsygv through SciPy.  Sizes are those the oracle assembles in seconds
(the small cells and the Si2 primitive cell).
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import oracle  # noqa: E402

SMALL_THRESH = 1.0e-8     # constants.f90:49 smallThresh


def int_consts(system):
    """intgConsts (potential.f90:150-160): multiplicity * pi^(3/2) / alpha^(3/2)."""
    out = []
    for p in system.pot_types:
        for a in p.alphas:
            out.append(float(len(p.sites)) * math.pi ** 1.5 / a ** 1.5)
    return np.array(out)


def num_electrons(system):
    """All-electron count of a neutral cell: Z times multiplicity of every
    potential type (makeinput's NUM_ELECTRONS with no core states)."""
    return float(sum(p.nuc_charge * len(p.sites) for p in system.pot_types))


def dposvx(A, B):
    """solveDPOSVX (interfaces.F90:236-293): dposvx('E', 'U') with A's upper
    triangle; returns (solution, info)."""
    from scipy.linalg import lapack
    A = np.asfortranarray(np.triu(A))
    B = np.asfortranarray(np.atleast_2d(np.asarray(B, dtype=np.float64).T).T)
    res = lapack.dposvx(A, B, fact="E", lower=0)
    x, info = res[5], res[9]
    return np.asarray(x), int(info)


def wigner(rho):
    """wignerXC / wignerXCEnergy (potentialUpdate.F90:2835-2894): the Wigner
    interpolation's potential and energy density (Hartree) at densities rho."""
    r13 = np.cbrt(rho)
    f = 1.0 + 12.57 * r13
    pot = -r13 * (0.984 + (0.943656 + 8.8963 * r13) / (f * f))
    en = -0.738 * r13 * (1.0 + 0.959 / f)
    return pot, en


class ScfPotential:
    """makeSCFPot for one system: the set-up matrices (elecStat.f90, the
    exchange-correlation mesh of exchCorr.f90) and the mixing history that
    the reference keeps in module variables across iterations (xl0..yl2,
    potentialUpdate.F90:38-45)."""

    def __init__(self, system, mesh=None, feedback_level=2, relax_factor=0.2,
                 convg_test=1.0e-4, electrons=None):
        self.system = system
        self.P = system.pot_dim
        self.feedback_level = feedback_level
        self.relax_factor = relax_factor
        self.convg_test = convg_test
        self.num_electrons = num_electrons(system) if electrons is None else electrons
        self.intg = int_consts(system)
        es = oracle.elecstat(system)
        self.pot_overlap = es["potAlphaOverlap"]
        self.non_local_neut = es["nonLocalNeutQPot"]
        self.local_neut = es["localNeutQPot"]
        self.non_local_nuc = es["nonLocalNucQPot"]
        self.local_nuc = es["localNucQPot"]
        self.residual_q = oracle.residual_q(system)
        self.xc = oracle.XcOracleMesh(system, mesh)
        self.xc_overlap = self.xc.overlap()
        # exchRhoOp and radialWeight of every site with a mesh, kept as the
        # reference keeps them in its HDF5 file and reads them back in every
        # makeSCFPot (potentialUpdate.F90:810-830)
        self.sites = []
        for i in range(len(system.pot_site_pos)):
            t = next(ti for ti, p in enumerate(system.pot_types) if i in p.sites)
            if system.pot_types[t].covalent_radius < SMALL_THRESH:
                continue
            n, rw, op = self.xc.site(i)
            self.sites.append((n, rw[:n].copy(), np.ascontiguousarray(op[:, :n, 0])))
        self.iteration = 1
        self.hist = None
        self.converged = False

    def xc_potential(self, rho_total, rho_core):
        """exchCorrPot(:, 1:3) before its solve (potentialUpdate.F90:833-1008,
        Wigner, spin 1, LDA) from the stored site operators: at every ray
        point the total and core densities op . rho (negative values set to
        0), then sum over points of radialWeight * (potential, energy
        density of the total, energy density of the core) * op."""
        out = np.zeros((self.P, 3))
        for n, rw, op in self.sites:
            tot = np.maximum(rho_total @ op, 0.0)
            core = np.maximum(rho_core @ op, 0.0)
            pot, en = wigner(tot)
            _, enc = wigner(core)
            out += op @ np.stack([rw * pot, rw * en, rw * enc], 1)
        return out

    def update(self, traces, pot_rho, pot_coeffs):
        """One makeSCFPot call.  traces = (chargeDensityTrace, nucPotTrace,
        kineticEnergyTrace); returns (new potCoeffs, record dict)."""
        P = self.P
        intg = self.intg
        sysm = self.system
        # charge fit: potAlphaOverlap x = [potRho, intgConsts]   (:191-239)
        g, info = dposvx(self.pot_overlap, np.stack([pot_rho, intg], 1))
        if info != 0:
            raise RuntimeError("Fitting dposvx failed. INFO= %d" % info)
        g1, g2 = g[:, 0].copy(), g[:, 1].copy()
        fitted = float(np.dot(intg, g1))
        electron_diff = self.num_electrons - fitted
        g2 = g1 + g2 * electron_diff / float(np.sum(intg * g2))      # :300-301
        g3 = np.zeros(P)                                              # no core (:331)
        g1 = g2 + g3
        g4 = g1.copy()
        g5 = np.zeros(P)
        g6 = np.zeros(P)
        g7 = np.zeros(P)
        # neutral charge per type (:416-484)
        off = 0
        for t, p in enumerate(sysm.pot_types):
            na = len(p.alphas)
            i0, i1 = off, off + na - 1
            nuc = p.nuc_charge * len(p.sites)
            s = float(np.sum(intg[i0:i1 + 1] * g1[i0:i1 + 1]))
            g5[i0] = (s - nuc) / intg[i0]
            g4[i0] = g1[i0] - g5[i0]
            g6[t] = g5[i0]
            g7[i1] = -nuc / intg[i1]
            off += na
        # electrostatic potential components (:536-606)
        e = np.zeros((P, 6))
        e[:, 0] = self.non_local_neut.T @ g4
        e[:, 1] = self.local_neut.T @ (g2 - g5)
        e[:, 2] = self.local_neut.T @ g3
        e[:, 3] = self.non_local_nuc
        e[:, 4] = self.local_nuc
        nt = len(sysm.pot_types)
        e[:, 5] = self.residual_q[:nt, :].T @ g6[:nt]
        es_energy = float(np.sum(0.5 * (g1 + g7) * (e[:, 0] + e[:, 5])
                                 + 0.5 * g2 * (e[:, 1] + e[:, 2] + e[:, 4])
                                 + 0.5 * (g2 - g5) * (e[:, 2] + e[:, 4])))
        es_energy = es_energy - float(np.sum(g2 * e[:, 3])) + traces[1]
        pot_es = np.stack([e[:, 0] + e[:, 1] + e[:, 2] - e[:, 3] + e[:, 4], e[:, 5]], 1)
        pot_es, info = dposvx(self.pot_overlap, pot_es)
        if info != 0:
            raise RuntimeError("Electrostatics dposvx failed. INFO= %d" % info)
        # exchange-correlation potential on the mesh (:720-1015)
        xc = self.xc_potential(g1, g3)                    # [P, 3]
        xc, info = dposvx(self.xc_overlap, xc)
        if info != 0:
            raise RuntimeError("Exchange-correlation dposvx failed. INFO= %d" % info)
        xc_energy = float(np.dot(self.pot_overlap.T @ xc[:, 1], g1)
                          - np.dot(self.pot_overlap.T @ xc[:, 2], g3))   # :1033-1039
        kinetic = float(traces[2])
        total = kinetic + es_energy + xc_energy
        if total > 0.0:
            raise RuntimeError("Total energy is positive (%.8f)" % total)
        # mixing history (:1301-1340)
        y_new = pot_es[:, 0] + pot_es[:, 1] + xc[:, 0]
        if self.hist is None:
            self.hist = {"xl1": pot_coeffs.copy(), "xl2": pot_coeffs.copy(),
                         "yl1": y_new.copy(), "yl2": y_new.copy()}
        h = self.hist
        xl0 = pot_coeffs.copy()
        yl0 = y_new
        out_pot = yl0 - xl0
        # RMS real-space potential change per site (:1350-1431)
        testable = 0.0
        for n, rw, op in self.sites:
            d = out_pot @ op                     # realSpacePotDiff(k) = op(:,k) . outputPot
            avg = math.sqrt(float(np.sum(rw * d * d)) / float(np.sum(rw)))
            testable = max(avg, testable)
        # Anderson-type feedback (:1604-1666)
        th1 = th2 = 0.0
        if self.feedback_level != 0 and self.iteration > 1:
            rl0 = yl0 - xl0
            rl1 = h["yl1"] - h["xl1"]
            rl2 = h["yl2"] - h["xl2"]
            d1 = rl0 - rl1
            d2 = rl0 - rl2
            S = self.pot_overlap
            u1 = S @ d1
            u2 = S @ d2
            s11 = float(np.dot(u1, d1))
            t1 = float(np.dot(u1, rl0))
            s12 = float(np.dot(u2, d1))
            s22 = float(np.dot(u2, d2))
            t2 = float(np.dot(u2, rl0))
            th1 = t1 / s11
            if self.feedback_level != 1 and self.iteration > 2:
                det = s11 * s22 - s12 * s12
                if det / (s11 * s22) >= 0.01:
                    th1 = (s22 * t1 - s12 * t2) / det
                    th2 = (-s12 * t1 + s11 * t2) / det
        r = self.relax_factor
        new = ((1.0 - r) * ((1.0 - th1 - th2) * xl0 + th1 * h["xl1"] + th2 * h["xl2"])
               + r * ((1.0 - th1 - th2) * yl0 + th1 * h["yl1"] + th2 * h["yl2"]))
        h["xl2"], h["xl1"] = h["xl1"], xl0
        h["yl2"], h["yl1"] = h["yl1"], yl0
        rec = dict(iteration=self.iteration, electron_diff=electron_diff,
                   testable_delta=testable, total_energy=total, kinetic_energy=kinetic,
                   elec_stat_energy=es_energy, exch_corr_energy=xc_energy,
                   rho_total=g1, rho_vale=g2)
        if testable < self.convg_test:
            self.converged = True
        else:
            self.iteration += 1
        return new, rec


def density_traces(system, vecs, occ, S, KE, NP, V):
    """makeValenceRho's traces for spin 1: per k-point rho(k) = sum_j f_j
    c_j c_j^H, then matrixElementMult with S, NP, KE and every V_j; sums
    over k.  vecs[k] is n x ns; S/KE/NP packed [nk][npack]; V [P][nk][npack]."""
    n = system.vale_dim
    cols = np.concatenate([np.full(c + 1, c) for c in range(n)])
    rows = np.concatenate([np.arange(c + 1) for c in range(n)])
    tr = np.zeros(3)
    prho = np.zeros(system.pot_dim)
    for k in range(system.num_kpoints):
        f = occ[k]
        if np.sum(np.abs(f)) <= SMALL_THRESH:
            continue
        X = vecs[k]
        rho = (X * f) @ X.conj().T
        rp = rho[rows, cols]
        tr[0] += oracle.matrix_element_mult(S[k], rp, n)
        tr[1] += oracle.matrix_element_mult(NP[k], rp, n)
        tr[2] += oracle.matrix_element_mult(KE[k], rp, n)
        for j in range(system.pot_dim):
            prho[j] += oracle.matrix_element_mult(V[j, k], rp, n)
    return tr, prho


def run_scf(system, coeffs0, max_iter=50, energy_tol=1.0e-6, mesh=None, num_states=None,
            feedback_level=2, relax_factor=0.2, convg_test=1.0e-4, log=None):
    """mainSCF on the oracle: returns (records, final coefficients).
    Converged when makeSCFPot's test holds (testableDelta < convgTest) and
    the total energy changed by less than energy_tol since the previous
    iteration."""
    n = system.vale_dim
    ns = n if num_states is None else num_states
    S = oracle.assemble(system, "ol")
    KE = oracle.assemble(system, "ke")
    NP = oracle.assemble(system, "np")
    V = oracle.assemble(system, "ep")
    Sf = [oracle.unpack(S[k], n) for k in range(system.num_kpoints)]
    pot = ScfPotential(system, mesh, feedback_level, relax_factor, convg_test)
    coeffs = np.array(coeffs0, dtype=np.float64)
    recs = []
    last_e = None
    for it in range(max_iter):
        evals = np.zeros((system.num_kpoints, ns))
        vecs = []
        for k in range(system.num_kpoints):
            H = oracle.build_hamiltonian(NP[k], KE[k], V[:, k], coeffs)
            w, v, info = oracle.hegv(oracle.unpack(H, n), Sf[k])
            if info != 0:
                raise RuntimeError("zhegv info %d at k-point %d" % (info, k + 1))
            evals[k] = w[:ns]
            vecs.append(v[:, :ns])
        occ = oracle.populate(evals, system.kpoint_weights, pot.num_electrons)
        tr, prho = density_traces(system, vecs, occ, S, KE, NP, V)
        coeffs, rec = pot.update(tr, prho, coeffs)
        rec["charge_trace"] = float(tr[0])
        de = None if last_e is None else rec["total_energy"] - last_e
        rec["energy_change"] = de
        last_e = rec["total_energy"]
        recs.append(rec)
        if log:
            log("oracle scf %d: E=%.10f dE=%s delta=%.3e" % (
                rec["iteration"], rec["total_energy"], "-" if de is None else "%.2e" % de,
                rec["testable_delta"]))
        if pot.converged and de is not None and abs(de) < energy_tol:
            break
        # the potential test holds but the energy still moves: go on; the
        # iteration number stays, as makeSCFPot leaves currIteration when
        # its test holds
        pot.converged = False
    return recs, coeffs
