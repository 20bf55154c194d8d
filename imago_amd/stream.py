"""The batched stream: many independent cells through the SCF on one GPU.

``stream_cells`` makes seeded cells of mixed C / Si / O (random sequential
insertion at 2.2 g/cm^3; at least 1.45 A between unlike atoms, 1.4 / 2.3 /
1.8 A between C / Si / O pairs), 2-128 atoms, with k-meshes of 1-8 irreducible points (the
smaller cells get the denser meshes) and basis orders from 28 to 1792.
``run`` feeds them to a pool of host threads, each with its own library
context and stream, so the integral set-up, the batched eigensolves and the
traces of several cells overlap on the device; the cooperative
tridiagonal reduction is taken by one context at a time.  Each cell
starts from the learned guess where a network exists (C, Si sites) and
the atomic superposition elsewhere, picks its eigensolve choices with the
autotuner (cached per order / k-points / scalar type), and runs the SCF
to 1e-6 Ha.  The report gives aggregate SCF solves per hour and the p50 /
p99 per-cell time to convergence.
"""
from __future__ import annotations

import argparse
import json
import math
import queue
import threading
import time

import numpy as np

from . import configs, guess, learned, lib, scf, tune
from .system import BOHR_PER_ANGSTROM, PotType, System, geometric_alphas, kpoint_mesh

MASS = {configs.SI: 28.0855, configs.O: 15.999, configs.C: 12.011}


def element_basis(element: int):
    if element in (configs.SI, configs.C):
        return learned.element_basis(element)
    rng = np.random.default_rng(101 + element)
    return configs.make_atom_type(rng, element, (0, 1, 2), nalpha=10, amin=0.15, amax=1.0e4)


# minimum distances (A): like pairs by species, unlike pairs 1.45
D_LIKE = {configs.C: 1.4, configs.SI: 2.3, configs.O: 1.8}


def _rsa(rng, box, species, d_like, d_unlike, max_tries=400000):
    pos, placed = [], []
    for sp in species:
        for _ in range(max_tries):
            p = rng.uniform(0, box, size=3)
            ok = True
            for q, sq in zip(pos, placed):
                d = p - q
                d -= box * np.round(d / box)
                if math.sqrt(float(d @ d)) < (d_like[sp] if sq == sp else d_unlike):
                    ok = False
                    break
            if ok:
                pos.append(p)
                placed.append(sp)
                break
        else:
            raise RuntimeError("random sequential insertion did not converge")
    return np.array(pos)


def stream_cell(rng, natoms, name):
    els = [configs.C, configs.SI, configs.O]
    frac = rng.dirichlet(np.ones(3))
    species = sorted(rng.choice(els, size=natoms, p=frac).tolist())
    mass = sum(MASS[s] for s in species)
    vol_a3 = mass / 6.02214076e23 / 2.2 * 1e24
    box_a = max(vol_a3 ** (1.0 / 3.0), 3.5)
    for _ in range(20):          # a jammed insertion retries 5 % larger
        try:
            pos = _rsa(rng, box_a, species, D_LIKE, 1.45, max_tries=20000) * BOHR_PER_ANGSTROM
            break
        except RuntimeError:
            box_a *= 1.05
    else:
        raise RuntimeError("no packing for %s" % name)
    box = box_a * BOHR_PER_ANGSTROM
    present = sorted(set(species))
    types = [element_basis(e) for e in present]
    tof = np.array([present.index(s) for s in species], dtype=np.int32)
    pots = [PotType(element=int(s), alphas=geometric_alphas(0.1, 1.0e4, 16),
                    nuc_charge=configs.Z_OF[int(s)], nuc_alpha=configs.NUC_ALPHA[int(s)],
                    sites=[i], covalent_radius=configs.COV_RADIUS[int(s)])
            for i, s in enumerate(species)]
    rv = np.eye(3) * box
    if natoms <= 8:
        mesh, gamma = (2, 2, 4), False          # 8 irreducible points
    elif natoms <= 48:
        mesh, gamma = (2, 2, 2), False          # 4
    else:
        mesh, gamma = (1, 1, 1), True           # Gamma
    kp, kw = kpoint_mesh(rv, mesh) if not gamma else (np.zeros((1, 3)), np.array([2.0]))
    amin = min(float(t.alphas[0]) for t in types)
    return System(real_vectors=rv, atom_types=types, atom_type_of=tof, atom_pos=pos,
                  pot_types=pots, pot_site_pos=pos.copy(), kpoints=kp, kpoint_weights=kw,
                  log_basis_fn_thresh=configs.log_thresh_for_cutoff(16.0, amin), gamma=gamma,
                  name=name)


def stream_cells(count, seed=50, min_atoms=2, max_atoms=128):
    """count cells, sizes log-uniform in [min_atoms, max_atoms]."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        na = int(round(math.exp(rng.uniform(math.log(min_atoms), math.log(max_atoms)))))
        out.append(stream_cell(rng, na, "cell%03d_%da" % (q, na)))
    return out


def _worker(cells, results, model, tuner, max_iter, log, device):
    db = guess.AtomicPotentials()
    with lib.Context(device) as ctx:
        mlps = {}
        if model is not None:
            for el, net in model.nets.items():
                m = lib.Mlp(ctx, learned.WIDTHS)
                m.set_params(net["params"])
                mlps[el] = m
        while True:
            try:
                q, cell = cells.get_nowait()
            except queue.Empty:
                break
            rec = dict(name=cell.name, atoms=cell.num_atoms, n=cell.vale_dim,
                       nk=cell.num_kpoints)
            try:
                t0 = time.perf_counter()
                atomic = guess.superposition_guess(ctx, cell, db)
                t_atoms = time.perf_counter() - t0
                t0 = time.perf_counter()
                scf.setup(ctx, cell)
                c0 = atomic if model is None else model.predict(ctx, cell, atomic, mlps)
                if tuner is not None:
                    rec["tune"] = tuner.choose(ctx, cell, c0)
                res = scf.run(ctx, cell, c0, max_iter=max_iter, do_setup=False)
                rec.update(seconds=time.perf_counter() - t0 + res.setup_s,
                           iterations=res.iterations, converged=res.converged,
                           energy=res.total_energy, atom_db_s=t_atoms)
            except lib.ImagoError as e:
                rec.update(error=str(e), converged=False)
            results[q] = rec
            if log:
                log("%s: %d atoms n=%d nk=%d -> %s" % (
                    cell.name, cell.num_atoms, cell.vale_dim, cell.num_kpoints,
                    "%d it, %.2f s, E=%.6f" % (rec["iterations"], rec["seconds"], rec["energy"])
                    if "seconds" in rec else rec.get("error")))
        for m in mlps.values():
            m.close()


def run(cells, workers=4, model=None, tuner=None, max_iter=100, log=None, device=0):
    """Every cell through the SCF on a pool of worker threads; returns
    (per-cell records, report)."""
    q = queue.Queue()
    for i, c in enumerate(cells):
        q.put((i, c))
    results = [None] * len(cells)
    t0 = time.perf_counter()
    threads = [threading.Thread(target=_worker,
                                args=(q, results, model, tuner, max_iter, log, device))
               for _ in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    done = [r for r in results if r and r.get("converged")]
    secs = np.array([r["seconds"] for r in done]) if done else np.zeros(1)
    report = dict(cells=len(cells), converged=len(done), workers=workers,
                  wall_s=wall, solves_per_hour=len(done) / wall * 3600.0,
                  p50_s=float(np.percentile(secs, 50)), p99_s=float(np.percentile(secs, 99)),
                  mean_iterations=float(np.mean([r["iterations"] for r in done])) if done else 0)
    return results, report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--max-iter", type=int, default=100)
    ap.add_argument("--guess", choices=("learned", "superposition"), default="learned")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)
    log = lambda *a: print("[stream]", *a, flush=True)
    cells = stream_cells(args.cells)
    model = None
    if args.guess == "learned":
        model = learned.LearnedGuess.load()
    tuner = None if args.no_tune else tune.Autotuner()
    results, report = run(cells, args.workers, model, tuner, args.max_iter, log)
    report["tuned"] = tuner.cache if tuner else None
    log(json.dumps(report))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(dict(report=report, cells=results), f, indent=1, default=float)
    return report


if __name__ == "__main__":
    main()
