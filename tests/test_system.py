"""Host-side set-up (imago_amd/system.py, configs.py): lattice, k-point
folding, basis normalisation, potential-term order."""
import math

import numpy as np
import pytest

from imago_amd import configs
from imago_amd.system import (ANG_NORM, _merge_sort_order, geometric_alphas, kpoint_mesh,
                              cubic_point_group, make_lattice, onsite_radial_norm)


def test_lattice_sorted_and_complete():
    rv = np.array([[6.0, 0.4, 0.0], [0.0, 5.5, 0.3], [0.2, 0.0, 7.0]]).T
    negl = 150.0
    dims, sizes = make_lattice(rv, negl)
    assert np.all(np.diff(sizes) >= 0)
    assert sizes[0] == 0.0 and np.all(dims[0] == 0)
    assert np.allclose(np.sum(dims ** 2, 1), sizes)
    # every lattice point within sqrt(negl) is present
    have = {tuple(np.round(d, 9)) for d in dims}
    for i in range(-4, 5):
        for j in range(-4, 5):
            for k in range(-4, 5):
                R = rv @ np.array([i, j, k])
                if R @ R <= negl:
                    assert tuple(np.round(R, 9)) in have


def _lattice_by_loops(real_vectors, negl):
    """makeLattice and cellBorderCheck (lattice.f90:571-770) loop by loop,
    with mergeSort's bottom-up merges (sortSubs.f90:18-177) written out."""
    reps = [2 * math.ceil(math.sqrt(negl) / math.sqrt(np.sum(real_vectors[:, i] ** 2))) + 1
            for i in range(3)]
    shift = real_vectors.sum(axis=1)
    dims, sizes = [], []
    for i in range(1, reps[0] + 1):
        for j in range(1, reps[1] + 1):
            for k in range(1, reps[2] + 1):
                rr = (i - (1 + reps[0]) // 2, j - (1 + reps[1]) // 2, k - (1 + reps[2]) // 2)
                d = np.zeros(3)
                for l, r in enumerate(rr):
                    for m in range(3):
                        d[m] = d[m] + real_vectors[m, l] * r
                sz = float(np.sum(d ** 2))
                if sz > negl:
                    t = d.copy()
                    border = False
                    for ax, sg in [(0, 1), (1, 1), (0, -1), (2, 1), (0, 1), (1, -1), (0, -1)]:
                        t[ax] += sg * shift[ax]
                        if np.sum(t * t) < negl:
                            border = True
                            break
                    if not border:
                        continue
                dims.append(d)
                sizes.append(sz)
    runs = [[q] for q in range(len(sizes))]
    while len(runs) > 1:
        merged = []
        for r in range(0, len(runs) - 1, 2):
            a, b = runs[r], runs[r + 1]
            ia = ib = 0
            out = []
            while ia < len(a) and ib < len(b):
                if sizes[a[ia]] < sizes[b[ib]]:
                    out.append(a[ia])
                    ia += 1
                else:
                    out.append(b[ib])
                    ib += 1
            merged.append(out + a[ia:] + b[ib:])
        if len(runs) % 2:
            merged.append(runs[-1])
        runs = merged
    order = runs[0]
    return np.array(dims)[order], np.array(sizes)[order]


@pytest.mark.parametrize("case", ["skewed", "si2_real", "si64_recip"])
def test_lattice_equals_the_loop_restatement_bit_for_bit(case):
    from imago_amd.system import reciprocal_vectors
    if case == "skewed":
        rv, negl = np.array([[6.0, 0.4, 0.0], [0.0, 5.5, 0.3], [0.2, 0.0, 7.0]]).T, 150.0
    elif case == "si2_real":
        s = configs.si2_primitive(mesh=(1, 1, 1))
        rv, negl = s.real_vectors, 400.0
    else:
        s = configs.si64_supercell(mesh=(1, 1, 1))
        rv, negl = reciprocal_vectors(s.real_vectors)[0], 4.0 * s.log_elec_thresh * s.min_pot_alpha
    d1, s1 = make_lattice(rv, negl)
    d2, s2 = _lattice_by_loops(rv, negl)
    assert np.array_equal(d1, d2) and np.array_equal(s1, s2)


def test_merge_sort_ties_take_second_run_first():
    # mergeLists compares with a strict '<': on equal keys the element of
    # the second run is emitted first (sortSubs.f90:120-177)
    order = _merge_sort_order(np.array([1.0, 1.0, 0.5, 1.0]))
    assert list(order) == [2, 3, 1, 0]


def test_si2_mesh_folds_to_ten_points():
    s = configs.si2_primitive()
    assert s.num_kpoints == 10
    assert s.kpoint_weights.sum() == pytest.approx(2.0)


def test_si64_mesh_has_four_time_reversal_pairs():
    s = configs.si64_supercell()
    assert s.num_kpoints == 4
    assert np.allclose(s.kpoint_weights, 0.5)
    assert s.vale_dim == 896 and s.pot_dim == 1024


def test_kpoint_folding_keeps_weight_and_symmetry():
    rv = 5.0 * np.eye(3)
    k, w = kpoint_mesh(rv, (4, 4, 4), ops=cubic_point_group())
    assert w.sum() == pytest.approx(2.0)
    k1, w1 = kpoint_mesh(rv, (4, 4, 4))
    assert len(k1) == 32 and np.allclose(w1, 2.0 / 32)


def test_geometric_alphas():
    a = geometric_alphas(0.12, 2.0e4, 12)
    assert a[0] == pytest.approx(0.12) and a[-1] == pytest.approx(2.0e4)
    assert np.allclose(a[1:] / a[:-1], a[1] / a[0])


def test_contracted_functions_are_normalised_on_site(oracle_mod):
    """<phi|phi> = 1 about one centre using the oracle's own overlap, for
    s, p and d shells of the Si64 basis (basis.f90 angNorm scaling)."""
    s = configs.si64_supercell()
    t = s.atom_types[0]
    z = np.zeros(3)
    al = t.alphas
    for si, (lm, col) in enumerate(t.states()):
        val = 0.0
        for a in range(len(al)):
            for b in range(len(al)):
                sh = oracle_mod.primitive("ol", (al[a], al[b], 0), z, z, z, 4 + 16 * 4)
                val += col[a] * col[b] * sh[lm - 1, lm - 1]
        assert val == pytest.approx(1.0, rel=1e-12), (si, lm, val)


def test_onsite_norm_formula_against_quadrature():
    al = np.array([0.5, 1.5])
    c = np.array([0.7, -0.3])
    r = np.linspace(0, 12, 200001)
    for l in range(3):
        f = (c[:, None] * np.exp(-al[:, None] * r ** 2)).sum(0) * r ** l
        num = np.trapezoid(f * f * r * r, r)
        assert num == pytest.approx(onsite_radial_norm(al, c, l), rel=1e-8)


def test_pot_terms_order_by_first_site():
    s = configs.small_test_system(seed=5)
    terms = s.pot_terms()
    firsts = [s.pot_types[t].sites[0] for t, _ in terms]
    assert firsts == sorted(firsts)
    assert len(terms) == s.pot_dim


def test_angnorm_are_unit_real_harmonics():
    # angNorm(lm)^2 * <combo^2>_sphere = 1 for every component
    rng = np.random.default_rng(0)
    v = rng.normal(size=(400000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    x, y, z = v.T
    combos = [np.ones_like(x), x, y, z, x * y, x * z, y * z, x * x - y * y, 2 * z * z - x * x - y * y,
              x * y * z, x * x * z - y * y * z, x ** 3 - 3 * y * y * x, 3 * x * x * y - y ** 3,
              2 * z ** 3 - 3 * x * x * z - 3 * y * y * z, 4 * z * z * x - x ** 3 - y * y * x,
              4 * z * z * y - x * x * y - y ** 3]
    for i, cmb in enumerate(combos):
        mean = np.mean((ANG_NORM[i] * cmb) ** 2) * 4 * math.pi
        assert mean == pytest.approx(1.0, rel=2e-2)


def test_sio2_builds_with_required_distances():
    s = configs.sio2_amorphous()
    assert s.num_atoms == 216 and s.vale_dim == 1944 and s.pot_dim == 3456
    box = s.real_vectors[0, 0]
    pos = s.atom_pos
    species = np.array([s.atom_types[t].element for t in s.atom_type_of])
    d = pos[:, None, :] - pos[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1)) / 1.8897261246257702
    np.fill_diagonal(r, 99)
    same = species[:, None] == species[None, :]
    assert r[same].min() >= 2.3 - 1e-9
    assert r[~same].min() >= 1.45 - 1e-9
