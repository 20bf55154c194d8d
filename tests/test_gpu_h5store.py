"""Checkpoint and resume of the integral stage on the GPU: matrices the
device assembled, written to the reference's SCF integral file and read
back into a fresh system, give the same panels bit for bit and the same
eigenvalues; the file holds the oracle's packed matrices; imago_set_packed
is the exact inverse of imago_get_packed."""
import numpy as np
import pytest

from conftest import rel_err
from imago_amd import configs, h5store
from imago_amd.lib import ImagoError

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not __import__("os").path.exists(h5store.H5LIB),
                                 reason="libimago_h5.so not built")]

KINDS = ("ol", "ke", "np", "ep")


def _raw(ctx, s, kind):
    terms = range(s.pot_dim) if kind == "ep" else [0]
    return [ctx.get_packed_raw(kind, t, k) for t in terms for k in range(s.num_kpoints)]


@pytest.mark.parametrize("gamma", [False, True])
def test_resume_reads_what_was_assembled(gpu_ctx, oracle_mod, tmp_path, gamma):
    s = configs.small_test_system(seed=11, gamma=gamma)
    mesh = (1, 1, 1) if gamma else (2, 2, 1)
    path = str(tmp_path / h5store.file_name())
    print("[gpu] h5 resume gamma=%s: first run assembles" % gamma, flush=True)
    gpu_ctx.set_system(s)
    with h5store.IntegralFile.for_system(path, s, mesh) as f:
        assert f.created
        assert h5store.assemble_or_load(gpu_ctx, f) == {k: "assembled" for k in KINDS}
        assert all(f.status(k, t) == 1 for k in KINDS
                   for t in (range(s.pot_dim) if k == "ep" else [0]))
        # the file holds the reference's packed matrices
        for kind in KINDS:
            ref = oracle_mod.assemble(s, kind)
            for k in range(s.num_kpoints):
                got = f.read(kind, 0, k)
                if not gamma:
                    got = got[0::2] + 1j * got[1::2]
                want = ref[0, k] if kind == "ep" else ref[k]
                assert rel_err(got, want) < 1e-10, kind
    panels = {kind: _raw(gpu_ctx, s, kind) for kind in KINDS}
    coeffs = configs.pot_coefficients(s, seed=4)
    ns = s.vale_dim - 1
    ev_a = gpu_ctx.secular(coeffs, ns)

    print("[gpu] h5 resume gamma=%s: second run loads" % gamma, flush=True)
    gpu_ctx.set_system(s)
    with h5store.IntegralFile.for_system(path, s, mesh) as f:
        assert not f.created
        assert h5store.assemble_or_load(gpu_ctx, f) == {k: "loaded" for k in KINDS}
    for kind in KINDS:
        for a, b in zip(panels[kind], _raw(gpu_ctx, s, kind)):
            assert np.array_equal(a, b), kind
    ev_b = gpu_ctx.secular(coeffs, ns)
    assert np.max(np.abs(ev_a - ev_b)) < 1e-12


def test_set_packed_inverts_get_packed(gpu_ctx):
    s = configs.small_test_system(seed=12)
    print("[gpu] set_packed round trip", flush=True)
    gpu_ctx.set_system(s)
    gpu_ctx.assemble("ke")
    rng = np.random.default_rng(5)
    for k in range(s.num_kpoints):
        p = gpu_ctx.get_packed_raw("ke", 0, k)
        x = np.where(p != 0.0, rng.standard_normal(p.size), 0.0)
        gpu_ctx.set_packed("ke", x, 0, k)
        assert np.array_equal(gpu_ctx.get_packed_raw("ke", 0, k), x)
    # complex values are accepted as such
    z = gpu_ctx.get_packed("ke", 0, 1) * (0.5 - 0.25j)
    gpu_ctx.set_packed("ke", z, 0, 1)
    assert np.array_equal(gpu_ctx.get_packed("ke", 0, 1), z)


def test_partially_loaded_kind_is_not_assembled(gpu_ctx):
    s = configs.small_test_system(seed=13)
    print("[gpu] partial load refused", flush=True)
    gpu_ctx.set_system(s)
    for kind in ("ke", "np", "ep"):
        gpu_ctx.assemble(kind)
    n = s.vale_dim
    one = np.zeros(2 * (n * (n + 1) // 2))
    one[0::2][[i * (i + 1) // 2 + i for i in range(n)]] = 1.0   # identity
    gpu_ctx.set_packed("ol", one, 0, 0)
    with pytest.raises(ImagoError):
        gpu_ctx.secular(configs.pot_coefficients(s), n)
    with pytest.raises(ImagoError):
        gpu_ctx.get_packed("ol", 0, 0)
    with pytest.raises(ImagoError):
        gpu_ctx.set_packed("ol", one, 0, s.num_kpoints)
