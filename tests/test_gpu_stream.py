"""The batched stream and the autotuner on the GPU: several small mixed
C / Si / O cells through the SCF on two worker threads (two contexts, two
streams) give the same converged energies as one context taking them one
after another, and every autotuner choice gives the same eigenvalues as
the default to 1e-10 Ha."""
import numpy as np
import pytest

from imago_amd import guess, lib, scf, stream, tune

pytestmark = pytest.mark.gpu


def test_autotuner_choices_agree(gpu_ctx):
    cells = stream.stream_cells(3, seed=7, min_atoms=4, max_atoms=12)
    tuner = tune.Autotuner()
    for cell in cells:
        print("[gpu] tune %s n=%d nk=%d" % (cell.name, cell.vale_dim, cell.num_kpoints), flush=True)
        db = guess.AtomicPotentials()
        c0 = guess.superposition_guess(gpu_ctx, cell, db)
        scf.setup(gpu_ctx, cell)
        ref = None
        for cand in tuner.candidates(cell):
            gpu_ctx.tune(**cand)
            ev = gpu_ctx.secular(c0, cell.vale_dim)
            if ref is None:
                ref = ev
            assert np.max(np.abs(ev - ref)) < 1e-10, cand
        choice = tuner.choose(gpu_ctx, cell, c0)
        assert choice in tuner.candidates(cell)
    gpu_ctx.tune(potrf_block=32, eigensolver=0, cholesky=0)


def test_stream_matches_one_context(gpu_ctx):
    cells = stream.stream_cells(6, seed=11, min_atoms=2, max_atoms=16)
    print("[gpu] stream of %d cells on 2 workers" % len(cells), flush=True)
    results, report = stream.run(cells, workers=2, model=None, tuner=tune.Autotuner(),
                                 max_iter=80)
    assert report["cells"] == 6
    for rec in results:
        if rec and rec.get("error"):
            print("[gpu] %s failed: %s" % (rec["name"], rec["error"]), flush=True)
    db = guess.AtomicPotentials()
    for cell, rec in zip(cells, results):
        if not rec.get("converged"):
            continue
        c0 = guess.superposition_guess(gpu_ctx, cell, db)
        scf.setup(gpu_ctx, cell)
        gpu_ctx.tune(**rec["tune"])
        res = scf.run(gpu_ctx, cell, c0, max_iter=80, do_setup=False)
        print("[gpu] %s: stream E %.10f, one context E %.10f" % (
            cell.name, rec["energy"], res.total_energy), flush=True)
        assert res.iterations == rec["iterations"]
        assert abs(res.total_energy - rec["energy"]) < 1e-8
    gpu_ctx.tune(potrf_block=32, eigensolver=0, cholesky=0)
    assert report["converged"] >= 4
