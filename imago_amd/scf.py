"""The SCF loop over the library: mainSCF (src/imago/imago.F90:488-523).

Each iteration is the reference's sequence

    secularEqnSCF   -> Context.secular         (H(k) build + batched zhegv)
    populateStates  -> lib.populate            (populateStandard)
    makeValenceRho  -> Context.density_traces  (rho(k), Tr(rho M), potRho)
    makeSCFPot      -> Context.scf_update      (fit, potential, energy, mixing)

until makeSCFPot's test holds (testableDelta < convgTest,
potentialUpdate.F90:1839) and, as the benchmark's criterion adds, the
total energy moved by less than ``energy_tol`` (1e-6 Ha) since the
previous iteration; at most LAST_ITERATION iterations (makeinputrc's 50).

``fort7_row`` writes the reference's iteration row (potentialUpdate.F90:
1806-1812, format ``(i5,3(1x,f13.8),3(1x,f17.8),1x,i3)``) and
``harvest_iteration_file`` restates the driver's read-out of it
(``_harvest_result``, src/scripts/imago.py:2630-2706): converged iff column
4 is below CONVERGENCE_TEST, total energy from column 5, the iteration
count from column 1, the gap (Hartree, column 7) in eV and its kind.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import lib

HARTREE_EV = 27.21138386                    # src/scripts/imago.py:2347
GAP_KIND_BY_CODE = {0: "none", 1: "direct", 2: "indirect"}


def num_electrons(system) -> float:
    """NUM_ELECTRONS of a neutral all-electron cell (no core states): Z
    times multiplicity of every potential type."""
    return float(sum(p.nuc_charge * len(p.sites) for p in system.pot_types))


@dataclass
class ScfResult:
    records: list = field(default_factory=list)   # one dict per iteration
    coeffs: np.ndarray = None                      # last potential coefficients
    rho: np.ndarray = None                         # fitted charge [2, potDim]
    converged: bool = False
    iterations: int = 0
    setup_s: float = 0.0
    loop_s: float = 0.0
    total_s: float = 0.0
    first_inputs: tuple = None    # (traces3, potRho, coefficients) of iteration 1

    @property
    def total_energy(self) -> float:
        return self.records[-1]["total_energy"] if self.records else float("nan")


PANEL_BUDGET_BYTES = 16 << 30


def ep_kind_for(system) -> str:
    """The EP path for a system: the direct path (no potDim panels, the H
    sum and the traces from the stored primitive blocks, four k-points per
    pass) for Gamma, up to 4 k-points, or when the potDim panels would take
    more than 16 GiB of HBM; the stored panels (one pass over them per
    iteration whatever nk) otherwise."""
    n = system.vale_dim
    panel = system.pot_dim * system.num_kpoints * n * (n + 1) // 2 * 16
    if system.gamma or system.num_kpoints <= 4 or panel > PANEL_BUDGET_BYTES:
        return "ep_direct"
    return "ep"


def setup(ctx, system, mesh=None, ep_kind=None, assemble=True, **params):
    """setupSCF's integral stage and makeSCFPot's set-up on the device."""
    t0 = time.perf_counter()
    ctx.set_system(system)
    if assemble:
        for kind in ("ol", "ke", "np", ep_kind or ep_kind_for(system)):
            ctx.assemble(kind)
    ctx.scf_setup(params.pop("electrons", num_electrons(system)), mesh=mesh, **params)
    return time.perf_counter() - t0


def run(ctx, system, coeffs0, max_iter=50, energy_tol=1.0e-6, mesh=None, num_states=None,
        ep_kind=None, log=None, do_setup=True, **params) -> ScfResult:
    """mainSCF on the GPU from the initial coefficients coeffs0 (the
    contents of scfV.dat, initPotCoeffs potential.f90:433-564)."""
    res = ScfResult()
    t_start = time.perf_counter()
    electrons = params.pop("electrons", num_electrons(system))
    if do_setup:
        res.setup_s = setup(ctx, system, mesh=mesh, ep_kind=ep_kind, electrons=electrons,
                            **params)
    n = system.vale_dim
    ns = n if num_states is None else num_states
    coeffs = np.array(coeffs0, dtype=np.float64, copy=True)
    last_e = None
    t_loop = time.perf_counter()
    for it in range(max_iter):
        t0 = time.perf_counter()
        ev = ctx.secular(coeffs, ns)
        t1 = time.perf_counter()
        occ, (e_occ, gap, gap_kind) = lib.populate(ev, system.kpoint_weights, electrons)
        t2 = time.perf_counter()
        tr, prho = ctx.density_traces(occ)
        t3 = time.perf_counter()
        if res.first_inputs is None:
            res.first_inputs = (tr.copy(), prho.copy(), coeffs.copy())
        coeffs, rec, rho = ctx.scf_update(tr, prho, coeffs)
        t4 = time.perf_counter()
        rec.update(occupied_energy=e_occ, band_gap=gap, gap_kind=gap_kind,
                   charge_trace=float(tr[0]), wall_s=t4 - t0,
                   stage_s=dict(secular=t1 - t0, populate=t2 - t1, traces=t3 - t2,
                                make_scf_pot=t4 - t3))
        de = None if last_e is None else rec["total_energy"] - last_e
        rec["energy_change"] = de
        last_e = rec["total_energy"]
        res.records.append(rec)
        res.rho = rho
        if log:
            log("scf %d: E=%.10f dE=%s delta=%.3e (%.1f ms)" % (
                rec["iteration"], rec["total_energy"], "-" if de is None else "%.2e" % de,
                rec["testable_delta"], rec["wall_s"] * 1e3))
        if rec["converged"] and de is not None and abs(de) < energy_tol:
            res.converged = True
            break
    res.coeffs = coeffs
    res.iterations = len(res.records)
    res.loop_s = time.perf_counter() - t_loop
    res.total_s = time.perf_counter() - t_start
    return res


def fort7_header() -> str:
    """The header line mainSCF writes to unit 7 (imago.F90:456-458)."""
    return (" Iter.  Occ._Energy     El._Error   Convergence    Total_Energy(au)      "
            "Mag._Mom.          Gap(au)  Gap_Kind")


def fort7_row(rec) -> str:
    """One iteration row, format (i5,3(1x,f13.8),3(1x,f17.8),1x,i3)
    (potentialUpdate.F90:1806-1808, spin 1: magnetic moment 0)."""
    return "%5d %13.8f %13.8f %13.8f %17.8f %17.8f %17.8f %3d" % (
        rec["iteration"], rec["occupied_energy"], rec["electron_diff"],
        rec["testable_delta"], rec["total_energy"], 0.0, rec["band_gap"], rec["gap_kind"])


def write_iteration_file(path, records):
    with open(path, "w") as f:
        f.write(fort7_header() + "\n")
        for r in records:
            f.write(fort7_row(r) + "\n")


def harvest_iteration_file(path, threshold):
    """The driver's read-out of the iteration file's last data row
    (_harvest_result, imago.py:2630-2706; _last_data_row 2599-2607):
    returns dict(converged, total_energy, iterations, magnetization,
    gap_ev, gap_kind); None fields when the file has no data row or the
    row is shorter than the column it would come from."""
    last = None
    with open(path) as f:
        for line in f:
            if line.strip():
                last = line
    out = dict(converged=False, total_energy=None, iterations=None, magnetization=None,
               gap_ev=None, gap_kind=None)
    row = last.split() if last is not None else None
    if row is None:
        return out
    out["converged"] = float(row[3]) < threshold
    out["total_energy"] = float(row[4])
    out["iterations"] = int(float(row[0]))
    if len(row) >= 6:
        out["magnetization"] = float(row[5])
    if len(row) >= 8:
        out["gap_ev"] = float(row[6]) * HARTREE_EV
        out["gap_kind"] = GAP_KIND_BY_CODE[int(float(row[7]))]
    return out
