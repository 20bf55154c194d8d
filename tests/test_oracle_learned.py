"""The oracle of the learned initial guess (oracle/mlp.py), pinned by
mathematics, not by restating it:

* the backpropagated gradient equals central finite differences of the loss;
* the descriptor is invariant under a rotation of the whole cell, a common
  translation and a permutation of the sites, equals its closed form for a
  single neighbour (l = 0: a; l = 1: a^2; l = 2: 2/3 a^2 with
  a = w g_k(r) f(r)), vanishes at the cut-off and is continuous across it;
* the first Adam step moves every parameter by lr * g / (|g| + eps).
"""
import numpy as np
import pytest

import mlp


def _cells(rv, n=2):
    r = np.arange(-n, n + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    return g @ rv.T


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = [7, 9, 6, 3]
    n = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(3))
    p = rng.normal(scale=0.5, size=n)
    x = rng.normal(size=(11, 7))
    t = rng.normal(size=(11, 3))
    loss, g = mlp.loss_grad(p, widths, x, t)
    h = 1e-6
    num = np.zeros(n)
    for q in range(n):
        e = np.zeros(n)
        e[q] = h
        num[q] = (mlp.loss_grad(p + e, widths, x, t)[0] - mlp.loss_grad(p - e, widths, x, t)[0]) / (2 * h)
    assert np.max(np.abs(num - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))
    assert loss == pytest.approx(np.mean((mlp.forward(p, widths, x) - t) ** 2))


def test_descriptor_invariances():
    rng = np.random.default_rng(1)
    rv = np.array([[7.0, 0.4, 0.0], [0.2, 6.5, 0.3], [0.0, 0.5, 7.2]])
    pos = rng.uniform(0, 7, size=(5, 3))
    w = np.array([1.0, 0.6, 1.0, 0.6, 1.0])
    args = (9.0, 0.5, 8.0)
    cells = _cells(rv)
    d0 = mlp.descriptor(pos, w, cells, *args)
    # rotation of the whole periodic system
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    d1 = mlp.descriptor(pos @ q.T, w, cells @ q.T, *args)
    assert np.allclose(d0, d1, rtol=1e-12, atol=1e-13)
    # common translation
    d2 = mlp.descriptor(pos + np.array([0.3, -1.2, 2.0]), w, cells, *args)
    assert np.allclose(d0, d2, rtol=1e-12, atol=1e-13)
    # permutation of the sites permutes the rows only
    perm = rng.permutation(5)
    d3 = mlp.descriptor(pos[perm], w[perm], cells, *args)
    assert np.allclose(d0[perm], d3, rtol=1e-12, atol=1e-13)


def test_descriptor_one_neighbour_closed_form():
    rc, rmin, eta = 9.0, 0.5, 3.0
    r = 3.7
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    w = np.array([1.0, 0.7])
    cells = np.zeros((1, 3))            # no periodic images
    d = mlp.descriptor(pos, w, cells, rc, rmin, eta)
    mu = rmin + np.arange(32) * (rc - rmin) / 31
    a = 0.7 * np.exp(-eta * (r - mu) ** 2) * 0.5 * (np.cos(np.pi * r / rc) + 1)
    assert np.allclose(d[0, :32], a, rtol=1e-14, atol=1e-300)
    assert np.allclose(d[0, 32:64], a * a, rtol=1e-13, atol=1e-300)
    assert np.allclose(d[0, 64:], 2.0 / 3.0 * a * a, rtol=1e-13, atol=1e-300)
    # the neighbour at the cut-off contributes nothing, and just inside it
    # almost nothing
    for rr in (rc, rc + 0.1):
        pos[1, 0] = rr
        assert np.all(mlp.descriptor(pos, w, cells, rc, rmin, eta)[0] == 0.0)
    pos[1, 0] = rc - 1e-6
    assert np.max(mlp.descriptor(pos, w, cells, rc, rmin, eta)[0]) < 1e-10


def test_first_adam_step_is_a_sign_step():
    rng = np.random.default_rng(3)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    lr, eps = 1e-2, 1e-8
    p1, m, v = mlp.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=lr, eps=eps)
    assert np.allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-15)
