"""Writes tests/golden/primitives.npz: golden primitive-integral blocks and
Boys values produced by the reference's own gaussIntegrals.f90, compiled
from the reference sources into oracle/_ref/libgauss_ref.so
(``make -C oracle ref``).

Inputs are seeded: for each of the 16 (l_a, l_b) classes (l1l2switch
codes of integrals.F90:533-536) and each kind (overlap2CIntg,
kinetic2CIntg, electron3CIntg, nuclear3CIntg) a handful of random tuples
with exponents spanning 0.1..2e4 and centres within a few bohr; Boys
values on a log grid plus points straddling the series/closed-form switch
at T = 10E-2.  Only the leading lmAlphaIndex block of each sh(16,16) is
meaningful; the whole block is stored (zeros elsewhere).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libgauss_ref.so not built (make -C oracle ref)")
    rng = np.random.default_rng(20261203)
    sws = [(1 << i) + (16 << j) for i in range(4) for j in range(4)]
    per_class = 4
    rec = {k: {"a": [], "xyz": [], "sw": [], "sh": []} for k in oracle.KINDS}
    for sw in sws:
        for t in range(per_class):
            a = np.exp(rng.uniform(np.log(0.1), np.log(2e4), 3))
            if t == 0:   # one diffuse tuple per class
                a = np.exp(rng.uniform(np.log(0.1), np.log(3.0), 3))
            A = rng.normal(scale=1.5, size=3)
            B = A + rng.normal(scale=1.0, size=3) / np.sqrt(a[:2].min())
            C = 0.5 * (A + B) + rng.normal(scale=1.0, size=3)
            # stay inside the region the assembly evaluates (alphaDist and
            # threeAlphaDist screens at the default -ln(1e-16))
            L = -np.log(1e-16)
            xi = a[0] * a[1] / (a[0] + a[1])
            d2 = np.sum((B - A) ** 2)
            if xi * d2 > L:
                B = A + (B - A) * np.sqrt(L / (xi * d2))
            z = a[0] + a[1]
            P = (a[0] * A + a[1] * B) / z
            w = z * a[2] / (z + a[2])
            c2 = np.sum((C - P) ** 2)
            if w * c2 > L:
                C = P + (C - P) * np.sqrt(L / (w * c2))
            for kind in oracle.KINDS:
                sh = oracle.primitive(kind, a, A, B, C, sw, source="ref")
                rec[kind]["a"].append(a)
                rec[kind]["xyz"].append(np.concatenate([A, B, C]))
                rec[kind]["sw"].append(sw)
                rec[kind]["sh"].append(np.array(sh))
    T = np.concatenate([np.exp(np.linspace(np.log(1e-8), np.log(80.0), 60)),
                        [0.0999999, 0.1, float(np.float32(0.1)), 0.1000001, 0.1001, 0.2, 0.5]])
    F = np.array([oracle.boys(t, source="ref") for t in T])
    out = {}
    for kind, r in rec.items():
        for key, v in r.items():
            out["%s_%s" % (kind, key)] = np.array(v)
    out["boys_T"] = T
    out["boys_F"] = F
    np.savez_compressed(os.path.join(HERE, "primitives.npz"), **out)
    print("wrote", os.path.join(HERE, "primitives.npz"))


if __name__ == "__main__":
    main()
