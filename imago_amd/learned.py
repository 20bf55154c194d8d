"""The learned initial SCF potential.

A per-site MLP (96 -> 128 -> 128 -> 16, FP32, trained with Adam by the
library's own kernels, imago_mlp_*) maps the environment descriptor of a
potential site (imago_site_descriptors: 32 radial channels x l = 0, 1, 2
within 6 A) to the correction of that site's 16 atomic-superposition
coefficients.  The guess is superposition + correction, written in the
order of the coefficient vector the reference reads from scfV.dat
(initPotCoeffs, src/imago/potential.f90:433-564); the superposition alone
stays the fallback (an element without a trained network, or a potential
type whose exponents differ from the network's).

Training data are converged SCF potentials the library itself produces on
seeded cells of C and Si (``training_cells``): diamond-like and
graphite-like (AB-stacked layered) lattices of 8-64 atoms, random strain
of up to 3 % per component and 0.08 A displacements, each converged from
the superposition guess.  ``python -m imago_amd.learned`` runs the
generation, training and held-out evaluation on the GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import configs, guess, lib, scf
from .system import (AtomType, PotType, System, BOHR_PER_ANGSTROM, geometric_alphas,
                     kpoint_mesh)

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "data", "learned_guess.npz")

CUTOFF = 6.0 * BOHR_PER_ANGSTROM        # descriptor cut-off, 6 A
R_MIN = 0.5                             # first radial centre (bohr)
ETA = 4.0                               # radial width exp(-eta (r - mu)^2), bohr^-2
WIDTHS = (96, 128, 128, 16)
NPOT = 16                               # potential terms per site


# -- workloads ---------------------------------------------------------------

def element_basis(element: int) -> AtomType:
    """One fixed basis per element: Si's is the Si64 workload's (s,s,p,s,p,d
    contracted from 12 exponents 0.12..2e4), C's is seeded the same way
    from 12 exponents 0.15..1.5e4."""
    if element == configs.SI:
        return configs.si64_supercell().atom_types[0]
    rng = np.random.default_rng(101 + element)
    return configs.make_atom_type(rng, element, (0, 0, 1, 0, 1, 2), nalpha=12, amin=0.15,
                                  amax=1.5e4)


def _cell_system(element, rv, frac, rng, strain, sigma_a, name, basis=None):
    eps = rng.uniform(-strain, strain, size=(3, 3))
    eps = 0.5 * (eps + eps.T)
    F = np.eye(3) + eps
    rv = F @ rv
    pos = frac @ rv.T
    pos = pos + rng.normal(scale=sigma_a * BOHR_PER_ANGSTROM, size=pos.shape)
    at = basis or element_basis(element)
    pots = [PotType(element=element, alphas=geometric_alphas(0.1, 1.0e4, NPOT),
                    nuc_charge=configs.Z_OF[element], nuc_alpha=configs.NUC_ALPHA[element],
                    sites=[i], covalent_radius=configs.COV_RADIUS[element])
            for i in range(len(pos))]
    kp, kw = kpoint_mesh(rv, (2, 2, 2))
    return System(real_vectors=rv, atom_types=[at], atom_type_of=np.zeros(len(pos), np.int32),
                  atom_pos=pos, pot_types=pots, pot_site_pos=pos.copy(), kpoints=kp,
                  kpoint_weights=kw,
                  log_basis_fn_thresh=configs.log_thresh_for_cutoff(16.0, float(at.alphas[0])),
                  name=name)


def diamond_cell(element, reps, rng, strain=0.03, sigma_a=0.08):
    a = (5.43 if element == configs.SI else 3.567) * BOHR_PER_ANGSTROM
    conv = configs.diamond_conventional(1.0)        # fractional
    frac = np.concatenate([(conv + np.array([i, j, k])) / np.array(reps, dtype=float)
                           for i in range(reps[0]) for j in range(reps[1])
                           for k in range(reps[2])])
    rv = np.diag([a * reps[0], a * reps[1], a * reps[2]])
    return _cell_system(element, rv, frac, rng, strain, sigma_a,
                        "diamond%d_%d" % (len(frac), element))


def layered_cell(element, reps, rng, strain=0.03, sigma_a=0.08):
    """AB-stacked hexagonal layers (graphite-like), 4 atoms per cell."""
    if element == configs.SI:
        a, c = 3.75, 2 * 3.9
    else:
        a, c = 2.46, 6.71
    a, c = a * BOHR_PER_ANGSTROM, c * BOHR_PER_ANGSTROM
    base = np.array([[0, 0, 0.25], [1 / 3, 2 / 3, 0.25], [0, 0, 0.75], [2 / 3, 1 / 3, 0.75]])
    frac = np.concatenate([(base + np.array([i, j, k])) / np.array(reps, dtype=float)
                           for i in range(reps[0]) for j in range(reps[1])
                           for k in range(reps[2])])
    a1 = np.array([a, 0, 0])
    a2 = np.array([-a / 2, a * math.sqrt(3) / 2, 0])
    a3 = np.array([0, 0, c])
    rv = np.stack([a1 * reps[0], a2 * reps[1], a3 * reps[2]], 1)
    return _cell_system(element, rv, frac, rng, strain, sigma_a,
                        "layered%d_%d" % (len(frac), element))


DIAMOND_REPS = {8: (1, 1, 1), 16: (2, 1, 1), 32: (2, 2, 1), 64: (2, 2, 2)}
LAYERED_REPS = {8: (2, 1, 1), 16: (2, 2, 1), 32: (2, 2, 2), 64: (4, 2, 2)}


def training_cells(count, seed, sizes=(8, 16, 32, 64)):
    """count seeded cells: element, lattice kind and size drawn in turn."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        el = (configs.C, configs.SI)[q % 2]
        kind = (q // 2) % 2
        size = sizes[(q // 4) % len(sizes)]
        if kind == 0:
            out.append(diamond_cell(el, DIAMOND_REPS[size], rng))
        else:
            out.append(layered_cell(el, LAYERED_REPS[size], rng))
    return out


# -- descriptors ---------------------------------------------------------------

def species_weight(element: int) -> float:
    return configs.Z_OF[element] / 14.0


def descriptor_cells(system: System, cutoff: float = CUTOFF) -> np.ndarray:
    """Every lattice translation that can bring a site image within the
    cut-off of another site of the cell."""
    rv = system.real_vectors
    pos = system.pot_site_pos
    span = np.max(np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1))
    recip = np.linalg.inv(rv)               # rows: reciprocal vectors / 2pi
    reach = cutoff + span
    n = [int(math.ceil(reach * np.linalg.norm(recip[i]))) for i in range(3)]
    g = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in n], indexing="ij"), -1)
    R = g.reshape(-1, 3) @ rv.T
    return R[np.linalg.norm(R, axis=1) <= reach + 1e-9]


def site_descriptors(ctx, system: System) -> np.ndarray:
    """[num_sites][96] on the GPU; each site weighted by its element as a
    neighbour."""
    w = np.zeros(len(system.pot_site_pos))
    for p in system.pot_types:
        for q in p.sites:
            w[q] = species_weight(p.element)
    return lib.site_descriptors(ctx, system.pot_site_pos, w, descriptor_cells(system), CUTOFF,
                                R_MIN, ETA)


def features(desc: np.ndarray) -> np.ndarray:
    """The network input: l = 0 channels as they are, the square roots of
    the l = 1 and l = 2 invariants (both are squares)."""
    f = desc.copy()
    f[:, 32:] = np.sqrt(np.maximum(f[:, 32:], 0.0))
    return f


# -- the model -----------------------------------------------------------------

class LearnedGuess:
    """Per-element networks with their input / output standardisation."""

    def __init__(self, nets=None):
        self.nets = nets or {}     # element -> dict(params, x_mean, x_std, y_mean, y_std, alphas)

    def save(self, path):
        arrs = {}
        for el, d in self.nets.items():
            for k, v in d.items():
                arrs["%d_%s" % (el, k)] = np.asarray(v)
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path=MODEL_PATH):
        z = np.load(path, allow_pickle=False)
        nets = {}
        for key in z.files:
            el, name = key.split("_", 1)
            nets.setdefault(int(el), {})[name] = z[key]
        return cls(nets)

    def predict(self, ctx, system: System, atomic: np.ndarray, mlps=None) -> np.ndarray:
        """Coefficients in the reference's term order: the superposition
        ``atomic`` plus each site's predicted correction."""
        desc = site_descriptors(ctx, system)
        f = features(desc)
        out = np.array(atomic, dtype=np.float64, copy=True)
        pos = 0
        term0 = {}
        for t, a in system.pot_terms():
            term0.setdefault(t, pos)
            pos += 1
        for el, net in self.nets.items():
            types = [t for t, p in enumerate(system.pot_types)
                     if p.element == el and len(p.sites) == 1 and len(p.alphas) == NPOT
                     and np.allclose(p.alphas, net["alphas"], rtol=1e-12)]
            if not types:
                continue
            sites = [system.pot_types[t].sites[0] for t in types]
            x = (f[sites] - net["x_mean"]) / net["x_std"]
            m = (mlps or {}).get(el)
            own = m is None
            if own:
                m = lib.Mlp(ctx, WIDTHS)
                m.set_params(net["params"])
            y = m.forward(x.astype(np.float32)).astype(np.float64)
            if own:
                m.close()
            dy = y * net["y_std"] + net["y_mean"]
            # how the correction is applied, chosen on validation cells
            # (select_application): a scale and the number of most diffuse
            # terms corrected (the sharp terms keep their atomic values)
            scale = float(net["scale"]) if "scale" in net else 1.0
            nterm = int(net["terms"]) if "terms" in net else NPOT
            for t, d in zip(types, dy):
                out[term0[t]:term0[t] + nterm] += scale * d[:nterm]
        return out

    def with_application(self, element, scale, terms):
        nets = {el: dict(net) for el, net in self.nets.items()}
        nets[element]["scale"] = np.float64(scale)
        nets[element]["terms"] = np.int64(terms)
        return LearnedGuess(nets)


# -- generation, training, evaluation ------------------------------------------

def converge(ctx, system, db, max_iter, log=None):
    """The plain GPU method on one cell: superposition guess, SCF to 1e-6 Ha."""
    atomic = guess.superposition_guess(ctx, system, db)
    res = scf.run(ctx, system, atomic, max_iter=max_iter)
    return atomic, res


def site_targets(system, atomic, coeffs):
    """Per site (one potential type per site): converged - atomic."""
    d = np.asarray(coeffs) - np.asarray(atomic)
    rows = {}
    pos = 0
    for t, a in system.pot_terms():
        rows.setdefault(t, []).append(pos)
        pos += 1
    out = np.zeros((len(system.pot_site_pos), NPOT))
    for t, p in enumerate(system.pot_types):
        out[p.sites[0]] = d[rows[t]]
    return out


def train_models(ctx, samples, epochs=400, minibatch=256, lr=1e-3, seed=0, log=None):
    """samples: element -> (features [N][96], targets [N][16], alphas)."""
    nets = {}
    for el, (x, y, alphas) in samples.items():
        xm, xs = x.mean(0), x.std(0) + 1e-8
        ym, ys = y.mean(0), y.std(0) + 1e-8
        m = lib.Mlp(ctx, WIDTHS, seed=seed + el)
        t0 = time.perf_counter()
        losses = m.train(((x - xm) / xs).astype(np.float32), ((y - ym) / ys).astype(np.float32),
                         epochs=epochs, minibatch=minibatch, lr=lr, seed=seed)
        if log:
            log("element %d: %d sites, loss %.4f -> %.4f in %d epochs (%.1f s)"
                % (el, len(x), losses[0], losses[-1], epochs, time.perf_counter() - t0))
        nets[el] = dict(params=m.get_params(), x_mean=xm, x_std=xs, y_mean=ym, y_std=ys,
                        alphas=alphas, losses=losses)
        m.close()
    return LearnedGuess(nets)


APPLICATIONS = ((1.0, 16), (0.5, 16), (1.0, 8), (0.5, 8), (1.0, 12))


def select_application(ctx, model, cells, db, max_iter, log=None):
    """Per element, the (scale, diffuse terms) of APPLICATIONS with the
    fewest SCF iterations summed over validation cells of that element
    (a cell that does not converge counts max_iter); the superposition
    alone is kept when no application beats it."""
    chosen = {}
    for el in list(model.nets):
        mine = [c for c in cells if c.pot_types[0].element == el]
        if not mine:
            continue
        base = []
        atomics = []
        for c in mine:
            a = guess.superposition_guess(ctx, c, db)
            atomics.append(a)
            r = scf.run(ctx, c, a, max_iter=max_iter)
            base.append(r.iterations if r.converged else max_iter)
        best = (sum(base), None)
        scores = {"superposition": sum(base)}
        for scale, terms in APPLICATIONS:
            trial = model.with_application(el, scale, terms)
            tot = 0
            for c, a in zip(mine, atomics):
                r = scf.run(ctx, c, trial.predict(ctx, c, a), max_iter=max_iter)
                tot += r.iterations if r.converged else max_iter
            scores["scale %.1f, %d terms" % (scale, terms)] = tot
            if tot < best[0]:
                best = (tot, (scale, terms))
        if log:
            log("element %d validation iterations: %s" % (el, scores))
        chosen[el] = dict(scores=scores, choice=best[1])
        if best[1] is None:
            model = model.with_application(el, 0.0, 0)
        else:
            model = model.with_application(el, *best[1])
    return model, chosen


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--cells", type=int, default=512)
    ap.add_argument("--eval", type=int, default=64, help="held-out 64-atom cells")
    ap.add_argument("--epochs", type=int, default=1000)
    ap.add_argument("--max-iter", type=int, default=100,
                    help="iterations allowed to a training cell (cells that do not converge "
                         "are left out)")
    ap.add_argument("--eval-max-iter", type=int, default=200)
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--report", default=None)
    ap.add_argument("--seed", type=int, default=20)
    ap.add_argument("--validation", type=int, default=16,
                    help="32-atom validation cells for choosing how the correction is applied")
    ap.add_argument("--model", default=None,
                    help="skip generation and training: select and evaluate this model")
    args = ap.parse_args(argv)
    log = lambda *a: print("[learned]", *a, flush=True)
    report = {"cells": args.cells, "eval_cells": args.eval, "epochs": args.epochs}
    with lib.Context(0) as ctx:
        db = guess.AtomicPotentials()
        if args.model:
            model = LearnedGuess.load(args.model)
        else:
            model = generate_and_train(ctx, db, args, report, log)
        t0 = time.perf_counter()
        vcells = training_cells(args.validation, args.seed + 500, sizes=(32,))
        model, chosen = select_application(ctx, model, vcells, db, args.eval_max_iter, log)
        report["selection"] = {str(k): v for k, v in chosen.items()}
        report["selection_s"] = time.perf_counter() - t0
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        model.save(args.out)
        evaluate(ctx, model, db, args, report, log)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1, default=float)
    return report


def generate_and_train(ctx, db, args, report, log):
    """Converge the training cells from the superposition guess, collect
    (descriptor, converged - atomic) per site, train one network per
    element."""
    feats, targs, alphas = {}, {}, {}
    t_gen = time.perf_counter()
    skipped = 0
    for q, cell in enumerate(training_cells(args.cells, args.seed)):
        t0 = time.perf_counter()
        try:
            atomic, res = converge(ctx, cell, db, args.max_iter)
        except lib.ImagoError as e:
            skipped += 1
            log("cell %d %s: %s, skipped" % (q, cell.name, e))
            continue
        if not res.converged:
            skipped += 1
            log("cell %d %s: not converged in %d iterations, skipped" % (q, cell.name,
                                                                         res.iterations))
            continue
        f = features(site_descriptors(ctx, cell))
        y = site_targets(cell, atomic, res.coeffs)
        el = cell.pot_types[0].element
        feats.setdefault(el, []).append(f)
        targs.setdefault(el, []).append(y)
        alphas[el] = cell.pot_types[0].alphas
        log("cell %d %s: %d atoms, %d iterations, E = %.8f, %.2f s" % (
            q, cell.name, cell.num_atoms, res.iterations, res.total_energy,
            time.perf_counter() - t0))
    report["generation_s"] = time.perf_counter() - t_gen
    report["skipped"] = skipped
    samples = {el: (np.concatenate(feats[el]), np.concatenate(targs[el]), alphas[el])
               for el in feats}
    t0 = time.perf_counter()
    model = train_models(ctx, samples, epochs=args.epochs, log=log)
    report["training_s"] = time.perf_counter() - t0
    report["training_sites"] = {str(el): int(len(samples[el][0])) for el in samples}
    report["final_loss"] = {str(el): float(model.nets[el]["losses"][-1]) for el in model.nets}
    return model


def evaluate(ctx, model, db, args, report, log):
    """Held-out 64-atom cells: the SCF from the superposition and from the
    learned guess, iterations, time and final energies side by side."""
    # held-out 64-atom cells: superposition vs learned guess
    rows = []
    rng_cells = training_cells(args.eval, args.seed + 1000, sizes=(64,))
    for q, cell in enumerate(rng_cells):
        atomic = guess.superposition_guess(ctx, cell, db)
        learned = model.predict(ctx, cell, atomic)
        r0 = scf.run(ctx, cell, atomic, max_iter=args.eval_max_iter)
        r1 = scf.run(ctx, cell, learned, max_iter=args.eval_max_iter)
        # how far each guess starts from the converged potential
        err_sup = float(np.linalg.norm(atomic - r0.coeffs))
        err_learned = float(np.linalg.norm(learned - r0.coeffs))
        rows.append(dict(name=cell.name, it_sup=r0.iterations, it_learned=r1.iterations,
                         conv_sup=r0.converged, conv_learned=r1.converged,
                         e_sup=r0.total_energy, e_learned=r1.total_energy,
                         t_sup=r0.total_s, t_learned=r1.total_s,
                         start_err_sup=err_sup, start_err_learned=err_learned))
        log("eval %d %s: superposition %d it (%.2f s), learned %d it (%.2f s), dE %.2e, "
            "start error %.3g -> %.3g" % (
                q, cell.name, r0.iterations, r0.total_s, r1.iterations, r1.total_s,
                r1.total_energy - r0.total_energy, err_sup, err_learned))
    report["eval"] = rows
    both = [r for r in rows if r["conv_sup"] and r["conv_learned"]]
    report["eval_converged_by_both"] = len(both)
    if both:
        it0 = np.array([r["it_sup"] for r in both])
        it1 = np.array([r["it_learned"] for r in both])
        t0s = np.array([r["t_sup"] for r in both])
        t1s = np.array([r["t_learned"] for r in both])
        report["mean_iterations_superposition"] = float(it0.mean())
        report["mean_iterations_learned"] = float(it1.mean())
        report["iteration_reduction"] = float(1.0 - it1.sum() / it0.sum())
        report["time_reduction"] = float(1.0 - t1s.sum() / t0s.sum())
        report["max_abs_energy_difference"] = float(max(abs(r["e_learned"] - r["e_sup"])
                                                        for r in both))
        log("held-out (%d of %d converged by both): %.1f -> %.1f iterations (%.0f %% "
            "fewer), max |dE| %.2e Ha" % (
                len(both), len(rows), it0.mean(), it1.mean(),
                100 * report["iteration_reduction"], report["max_abs_energy_difference"]))
    cap = args.eval_max_iter
    it0 = np.array([r["it_sup"] if r["conv_sup"] else cap for r in rows])
    it1 = np.array([r["it_learned"] if r["conv_learned"] else cap for r in rows])
    report["all_cells_iterations"] = dict(superposition=float(it0.mean()),
                                          learned=float(it1.mean()),
                                          converged_superposition=int(sum(
                                              r["conv_sup"] for r in rows)),
                                          converged_learned=int(sum(
                                              r["conv_learned"] for r in rows)))


if __name__ == "__main__":
    main()
