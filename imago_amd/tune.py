"""The autotuner of the batched eigensolve.

imago_tune exposes the run-time choices of the per-iteration stage: the
columns per diagonal block of the blocked Cholesky (8, 16, 24, 32), and the
eigensolver (this library's cooperative tridiagonal reduction with divide
and conquer and the compact-WY back-transform, or rocSOLVER's heevd /
syevd, which the tuner offers only beyond the orders the library's
reduction takes).  The best choice depends on the matrix order, the batch (k-points)
and the scalar type; measured optima are 32 columns for 4 x 896 complex
and 16 for 1 x 1944 real.  ``Autotuner`` times each
candidate on the system the context holds, two passes of imago_secular,
populate and imago_density_traces per candidate (the first also factors the
overlap; the traces finish the occupied eigenvector columns the eigensolve
leaves pending), scores a candidate by the cost of an SCF of
``iterations`` iterations, keeps the fastest per
(order rounded up to 128, k-points, complex / real) in a cache (a JSON file when a path is
given) and applies it.
"""
from __future__ import annotations

import json
import os
import threading
import time

BLOCKS = (8, 16, 24, 32)


def key(system) -> str:
    """The class a choice is kept for: the order rounded up to a multiple
    of 128, the k-point count, complex or real."""
    n = -(-system.vale_dim // 128) * 128
    return "n%d_k%d_%s" % (n, system.num_kpoints, "real" if system.gamma else "cx")


def coop_supported(system) -> bool:
    """Orders the library's own tridiagonal reduction takes (1024 complex,
    2048 real)."""
    return system.vale_dim <= (2048 if system.gamma else 1024)


class Autotuner:
    def __init__(self, path: str | None = None, iterations: int = 30):
        self.path = path
        self.iterations = iterations
        self.cache = {}
        self.log = {}
        self._lock = threading.Lock()
        if path and os.path.exists(path):
            with open(path) as f:
                self.cache = json.load(f)

    def candidates(self, system):
        """Cholesky block sizes with this library's eigensolver, or with
        rocSOLVER's heevd / syevd only for orders the library's reduction
        does not take: rocSOLVER's factorizations have returned wrong
        results now and then while other contexts drove the device from
        other host threads, as the batched stream does."""
        solvers = (0,) if coop_supported(system) else (1,)
        return [dict(potrf_block=b, eigensolver=e) for e in solvers for b in BLOCKS]

    def choose(self, ctx, system, coeffs):
        """The cached choice for this system's class, or a timing of every
        candidate on it (the system must be set and assembled on ctx)."""
        k = key(system)
        with self._lock:
            got = self.cache.get(k)
        if got is None:
            from . import lib, scf
            n = system.vale_dim
            electrons = scf.num_electrons(system)

            def one_pass():
                ev = ctx.secular(coeffs, n)
                occ, _ = lib.populate(ev, system.kpoint_weights, electrons)
                ctx.density_traces(occ)

            scores = []
            for cand in self.candidates(system):
                ctx.tune(**cand)
                t0 = time.perf_counter()
                one_pass()
                t1 = time.perf_counter()
                one_pass()
                t2 = time.perf_counter()
                first, steady = t1 - t0, t2 - t1
                scores.append((first + (self.iterations - 1) * steady, cand, first, steady))
            best = min(scores, key=lambda x: x[0])
            got = dict(best[1])
            with self._lock:
                self.cache[k] = got
                self.log[k] = [dict(c, first_ms=round(f * 1e3, 3), steady_ms=round(s * 1e3, 3))
                               for _, c, f, s in scores]
                if self.path:
                    os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
                    with open(self.path, "w") as f:
                        json.dump(self.cache, f, indent=1)
        ctx.tune(**got)
        return got
