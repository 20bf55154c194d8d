"""NumPy restatement of the learned initial guess's kernels: test
infrastructure only (the checker of imago_site_descriptors and the imago_mlp
entry points; nothing in the product imports it).

* ``descriptor``   the 96 rotation invariants of each site's environment
                   (32 Gaussian radial channels x l = 0, 1, 2)
* ``forward``      the MLP (tanh hidden layers, linear output), parameters
                   packed per layer as W_l [w_{l+1}][w_l] row-major, b_l
* ``loss_grad``    mean squared error and its gradient by backpropagation
* ``adam``         one Adam step (Kingma and Ba 2015, bias-corrected)

The reference holds no learned guess (its SCF starts from scfV.dat,
src/imago/potential.f90:433-564); the functions are pinned by
mathematics in tests/test_oracle_learned.py: finite differences for the
gradient, rotation / translation / permutation invariance and a
one-neighbour closed form for the descriptor, the first Adam step's
closed form.
"""
from __future__ import annotations

import numpy as np

NRAD = 32


def descriptor(pos, weight, cells, cutoff, r_min, eta):
    """[num_sites][96] in float64."""
    pos = np.asarray(pos, dtype=np.float64)
    w = np.asarray(weight, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.float64)
    mu = r_min + np.arange(NRAD) * (cutoff - r_min) / (NRAD - 1)
    out = np.zeros((len(pos), 3 * NRAD))
    for i in range(len(pos)):
        d = (pos[None, :, :] - pos[i][None, None, :] + cells[:, None, :]).reshape(-1, 3)
        wj = np.tile(w, len(cells))
        r2 = np.einsum("ij,ij->i", d, d)
        keep = (r2 < cutoff * cutoff) & (r2 >= 1e-20)
        d, wj, r = d[keep], wj[keep], np.sqrt(r2[keep])
        u = d / r[:, None]
        fc = 0.5 * (np.cos(np.pi * r / cutoff) + 1.0)
        a = wj[:, None] * np.exp(-eta * (r[:, None] - mu[None, :]) ** 2) * fc[:, None]  # [nb, K]
        out[i, :NRAD] = a.sum(0)
        v = a.T @ u                                                   # [K, 3]
        out[i, NRAD:2 * NRAD] = np.einsum("kx,kx->k", v, v)
        q = np.einsum("nk,nx,ny->kxy", a, u, u) - a.sum(0)[:, None, None] * np.eye(3) / 3.0
        out[i, 2 * NRAD:] = np.einsum("kxy,kxy->k", q, q)
    return out


def unpack(params, widths):
    layers = []
    o = 0
    for l in range(len(widths) - 1):
        i, n = widths[l], widths[l + 1]
        W = params[o:o + n * i].reshape(n, i)
        o += n * i
        b = params[o:o + n]
        o += n
        layers.append((W, b))
    return layers


def forward(params, widths, x, keep=False):
    """MLP output [batch][w_L]; with keep, also the activations."""
    h = [np.asarray(x, dtype=np.float64)]
    layers = unpack(np.asarray(params, dtype=np.float64), widths)
    for l, (W, b) in enumerate(layers):
        z = h[-1] @ W.T + b
        h.append(np.tanh(z) if l + 1 < len(layers) else z)
    return (h[-1], h) if keep else h[-1]


def loss_grad(params, widths, x, target):
    """(mean squared error, gradient packed like params)."""
    y, h = forward(params, widths, x, keep=True)
    t = np.asarray(target, dtype=np.float64)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = 2.0 * (y - t) / n
    layers = unpack(np.asarray(params, dtype=np.float64), widths)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (dz.T @ h[l], dz.sum(0))
        if l > 0:
            dh = dz @ W
            dz = dh * (1.0 - h[l] ** 2)
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam(params, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step at step number ``step`` (1-based); returns (params, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
